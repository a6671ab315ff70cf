// spec.hip -- structured elasticity problems generated on the device from a
// compact spec (sfem_problem_spec): meshgrid coordinates and connectivity,
// per-element Young's moduli, the clamped face, and consistent nodal loads
// (uniform traction on a face, a body force, seeded random nodal loads).
//
// The reference builds its meshes and constraints on the host
// (meshgrid, examples/vector3D.jl:10-130; setboundaryCond, :133-173) and has
// no load assembly; a request loop that receives arrays spends its time
// preparing and copying them.  Here a request is ~150 bytes and everything
// is formed in HBM.
//
// Random numbers are counter-based, so a value depends only on (seed, stream,
// index): stream 1 draws the element moduli, stream 2 the nodal noise.
//   z = splitmix64 finaliser of (seed ^ stream*C1 ^ index*C2) + C2
//   u = ((z >> 12) + 1/2) 2^-52, exactly representable, in (0, 1)
//   E_e = exp(log lo + u_e (log hi - log lo))    (E = lo when lo == hi)
//   noise of local dof k = sigma sqrt(-2 log u(2k)) cos(2 pi u(2k+1))
// Consistent loads are integrated with the 2-point Gauss rule of the
// element (2x2x2 cells, 2x2 faces, 2-point edges), and each node sums its
// cells (faces) in ascending cell order, so the result is deterministic.
#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

// the load integrals are the exact sequence of roundings a host
// restatement repeats (no multiply-add contraction)
#pragma clang fp contract(off)

#include "basis.hpp"
#include "batch.hpp"

namespace sfem {
namespace {

constexpr uint64_t kC1 = 0xD1B54A32D192ED03ull, kC2 = 0x9E3779B97F4A7C15ull;

__device__ __forceinline__ double spec_u01(uint64_t seed, uint64_t stream, uint64_t index) {
    uint64_t z = (seed ^ (stream * kC1) ^ (index * kC2)) + kC2;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return ((double)(z >> 12) + 0.5) * 0x1p-52;
}

// One problem as the kernels see it (device copy of the spec plus offsets).
struct SpecDev {
    int ndim, clamp_axis, clamp_side, load_axis, load_side, pad;
    int64_t ne[3], nn[3];  // elements / nodes per direction (1 beyond ndim)
    double box[6];
    double ylo, yhi, nu, noise;
    double trac[3], body[3];
    uint64_t seed;
    int64_t node_off, elem_off;
};

// numpy / Julia linspace point i of ne intervals, end point exact (as meshgrid)
__device__ __forceinline__ double lin_pt(double x0, double x1, int64_t ne, int64_t i) {
    if (i == ne) return x1;
    return (double)i * ((x1 - x0) / (double)ne) + x0;
}

__device__ __forceinline__ int find_problem(const int64_t* __restrict__ off, int nprob, int64_t g) {
    int lo = 0, hi = nprob - 1;  // last p with off[p] <= g
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (off[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ double coord_of(const SpecDev& P, int d, int64_t i) {
    return lin_pt(P.box[2 * d], P.box[2 * d + 1], P.ne[d], i);
}

// int N_a dV over cell (cx, cy, cz) for its corner a (Q1 corner order), 2x2x2 Gauss
__device__ double cell_load_3d(const SpecDev& P, int64_t cx, int64_t cy, int64_t cz, int a) {
    double X[8][3];
    for (int b = 0; b < 8; ++b) {
        const int64_t ix = cx + (q1_sx(b) > 0), iy = cy + (q1_sy(b) > 0), iz = cz + (q1_sz(b) > 0);
        X[b][0] = coord_of(P, 0, ix);
        X[b][1] = coord_of(P, 1, iy);
        X[b][2] = coord_of(P, 2, iz);
    }
    double s = 0.0;
    for (int g = 0; g < 8; ++g) {
        double xi, eta, zeta;
        gauss_point(g, xi, eta, zeta);
        double J[3][3] = {};
        for (int b = 0; b < 8; ++b) {
            double dn[3];
            q1_3d_dN(b, xi, eta, zeta, dn);
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) J[i][j] = J[i][j] + X[b][i] * dn[j];
        }
        const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) +
                           J[0][1] * (J[1][2] * J[2][0] - J[1][0] * J[2][2]) +
                           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        s = s + q1_3d_N(a, xi, eta, zeta) * fabs(det);
    }
    return s;
}

// int N_a dA over cell (cx, cy) of a Q4 mesh, 2x2 Gauss
__device__ double cell_load_2d(const SpecDev& P, int64_t cx, int64_t cy, int a) {
    double X[4][2];
    for (int b = 0; b < 4; ++b) {
        X[b][0] = coord_of(P, 0, cx + (q1_sx(b) > 0));
        X[b][1] = coord_of(P, 1, cy + (q1_sy(b) > 0));
    }
    double s = 0.0;
    for (int g = 0; g < 4; ++g) {
        double xi, eta, zeta;
        gauss_point(g, xi, eta, zeta);
        double J[2][2] = {};
        for (int b = 0; b < 4; ++b) {
            double dn[2];
            q1_2d_dN(b, xi, eta, dn);
            for (int i = 0; i < 2; ++i)
                for (int j = 0; j < 2; ++j) J[i][j] = J[i][j] + X[b][i] * dn[j];
        }
        s = s + q1_2d_N(a, xi, eta) * fabs(J[0][0] * J[1][1] - J[0][1] * J[1][0]);
    }
    return s;
}

// int N_a dA over the face of the load plane with lower corner (iu, iv)
// along the plane's two other axes u < v (3D), 2x2 Gauss; a in the Q4
// corner order of (u, v)
__device__ double face_load_3d(const SpecDev& P, int du, int dv, int64_t iw, int64_t iu, int64_t iv, int a) {
    const int dw = P.load_axis;
    double X[4][3];
    for (int b = 0; b < 4; ++b) {
        int64_t idx[3];
        idx[dw] = iw;
        idx[du] = iu + (q1_sx(b) > 0);
        idx[dv] = iv + (q1_sy(b) > 0);
        for (int d = 0; d < 3; ++d) X[b][d] = coord_of(P, d, idx[d]);
    }
    double s = 0.0;
    for (int g = 0; g < 4; ++g) {
        double xi, eta, zeta;
        gauss_point(g, xi, eta, zeta);
        double t1[3] = {0, 0, 0}, t2[3] = {0, 0, 0};
        for (int b = 0; b < 4; ++b) {
            double dn[2];
            q1_2d_dN(b, xi, eta, dn);
            for (int d = 0; d < 3; ++d) {
                t1[d] = t1[d] + X[b][d] * dn[0];
                t2[d] = t2[d] + X[b][d] * dn[1];
            }
        }
        const double c0 = t1[1] * t2[2] - t1[2] * t2[1], c1 = t1[2] * t2[0] - t1[0] * t2[2],
                     c2 = t1[0] * t2[1] - t1[1] * t2[0];
        s = s + q1_2d_N(a, xi, eta) * sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    }
    return s;
}

// int N_a ds over the edge of the load line (2D) starting at index iu along
// the other axis, 2-point Gauss; a = 0 (lower end) or 1
__device__ double edge_load_2d(const SpecDev& P, int du, int64_t iw, int64_t iu, int a) {
    const int dw = P.load_axis;
    double X[2][2];
    for (int b = 0; b < 2; ++b) {
        int64_t idx[2];
        idx[dw] = iw;
        idx[du] = iu + b;
        for (int d = 0; d < 2; ++d) X[b][d] = coord_of(P, d, idx[d]);
    }
    const double gs = gauss2_abscissa();
    double s = 0.0;
    for (int g = 0; g < 2; ++g) {
        const double xi = g == 0 ? -gs : gs;
        const double t0 = (X[1][0] - X[0][0]) / 2, t1 = (X[1][1] - X[0][1]) / 2;
        const double N = a == 0 ? (1 - xi) / 2 : (1 + xi) / 2;
        s = s + N * sqrt(t0 * t0 + t1 * t1);
    }
    return s;
}

// One thread per node of the concatenated problems: coordinates, dof ids,
// the clamp (every dof of the clamped face), and the consistent loads.
__global__ __launch_bounds__(256) void spec_nodes_kernel(const SpecDev* __restrict__ specs, int nprob,
                                                         const int64_t* __restrict__ noff, int64_t n_total,
                                                         double* __restrict__ coords, int32_t* __restrict__ dof,
                                                         double* __restrict__ f, uint8_t* __restrict__ fixed) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_total) return;
    const SpecDev& P = specs[find_problem(noff, nprob, g)];
    const int nd = P.ndim;
    const int64_t m = g - P.node_off;
    int64_t idx[3];
    idx[0] = m % P.nn[0];
    idx[1] = (m / P.nn[0]) % P.nn[1];
    idx[2] = nd == 3 ? m / (P.nn[0] * P.nn[1]) : 0;
    for (int d = 0; d < nd; ++d) {
        coords[g * nd + d] = coord_of(P, d, idx[d]);
        dof[g * nd + d] = (int32_t)(nd * g + d);
    }
    const bool clamped = idx[P.clamp_axis] == (P.clamp_side ? P.nn[P.clamp_axis] - 1 : 0);
    const bool loaded = idx[P.load_axis] == (P.load_side ? P.nn[P.load_axis] - 1 : 0);
    double load[3] = {0.0, 0.0, 0.0};
    // body force: the cells around the node, ascending cell index
    if (P.body[0] != 0.0 || P.body[1] != 0.0 || (nd == 3 && P.body[2] != 0.0)) {
        double w = 0.0;
        for (int oz = nd == 3 ? -1 : 0; oz <= 0; ++oz)
            for (int oy = -1; oy <= 0; ++oy)
                for (int ox = -1; ox <= 0; ++ox) {
                    const int64_t cx = idx[0] + ox, cy = idx[1] + oy, cz = idx[2] + oz;
                    if (cx < 0 || cx >= P.ne[0] || cy < 0 || cy >= P.ne[1] || cz < 0 || cz >= P.ne[2]) continue;
                    // the node's corner in the cell: offsets (-ox, -oy, -oz) from the cell's lower corner
                    const int a = nd == 3 ? ((-ox) ^ (-oy)) + 2 * (-oy) + 4 * (-oz) : ((-ox) ^ (-oy)) + 2 * (-oy);
                    w = w + (nd == 3 ? cell_load_3d(P, cx, cy, cz, a) : cell_load_2d(P, cx, cy, a));
                }
        for (int d = 0; d < nd; ++d) load[d] = load[d] + P.body[d] * w;
    }
    // uniform traction on the loaded face: the faces (edges) around the node
    if (loaded && (P.trac[0] != 0.0 || P.trac[1] != 0.0 || (nd == 3 && P.trac[2] != 0.0))) {
        double w = 0.0;
        if (nd == 3) {
            const int du = P.load_axis == 0 ? 1 : 0, dv = P.load_axis == 2 ? 1 : 2;
            for (int ov = -1; ov <= 0; ++ov)
                for (int ou = -1; ou <= 0; ++ou) {
                    const int64_t iu = idx[du] + ou, iv = idx[dv] + ov;
                    if (iu < 0 || iu >= P.ne[du] || iv < 0 || iv >= P.ne[dv]) continue;
                    const int a = ((-ou) ^ (-ov)) + 2 * (-ov);
                    w = w + face_load_3d(P, du, dv, idx[P.load_axis], iu, iv, a);
                }
        } else {
            const int du = 1 - P.load_axis;
            for (int ou = -1; ou <= 0; ++ou) {
                const int64_t iu = idx[du] + ou;
                if (iu < 0 || iu >= P.ne[du]) continue;
                w = w + edge_load_2d(P, du, idx[P.load_axis], iu, -ou);
            }
        }
        for (int d = 0; d < nd; ++d) load[d] = load[d] + P.trac[d] * w;
    }
    for (int d = 0; d < nd; ++d) {
        const int64_t k = m * nd + d;  // local dof
        double v = load[d];
        if (loaded && !clamped && P.noise != 0.0) {
            const double u1 = spec_u01(P.seed, 2, 2 * (uint64_t)k), u2 = spec_u01(P.seed, 2, 2 * (uint64_t)k + 1);
            v = v + P.noise * (sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
        }
        f[g * nd + d] = v;
        fixed[g * nd + d] = clamped ? 1 : 0;
    }
}

// One thread per element: meshgrid connectivity (global node ids, 0-based,
// the reference's corner order), Young's modulus, Poisson's ratio.
__global__ __launch_bounds__(256) void spec_elems_kernel(const SpecDev* __restrict__ specs, int nprob,
                                                         const int64_t* __restrict__ eoff, int64_t e_total,
                                                         int32_t* __restrict__ ien, double* __restrict__ young,
                                                         double* __restrict__ nu) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= e_total) return;
    const SpecDev& P = specs[find_problem(eoff, nprob, g)];
    const int64_t e = g - P.elem_off;
    const int64_t nx = P.ne[0], ny = P.ne[1], row = nx + 1;
    if (P.ndim == 2) {
        const int64_t i = e % nx, j = e / nx;
        const int64_t n0 = P.node_off + j * row + i;
        int32_t* c = ien + g * 4;
        c[0] = (int32_t)n0;
        c[1] = (int32_t)(n0 + 1);
        c[2] = (int32_t)(n0 + row + 1);
        c[3] = (int32_t)(n0 + row);
    } else {
        const int64_t i = e % nx, j = (e / nx) % ny, k = e / (nx * ny);
        const int64_t p = (nx + 1) * (ny + 1);
        const int64_t n0 = P.node_off + k * p + j * row + i;
        int32_t* c = ien + g * 8;
        c[0] = (int32_t)n0;
        c[1] = (int32_t)(n0 + 1);
        c[2] = (int32_t)(n0 + row + 1);
        c[3] = (int32_t)(n0 + row);
        for (int a = 0; a < 4; ++a) c[4 + a] = (int32_t)(c[a] + p);
    }
    double E = P.ylo;
    if (P.yhi != P.ylo) {
        const double u = spec_u01(P.seed, 1, (uint64_t)e);
        E = exp(log(P.ylo) + u * (log(P.yhi) - log(P.ylo)));
    }
    young[g] = E;
    nu[g] = P.nu;
}

void check_spec(const sfem_problem_spec& s, int64_t index) {
    const std::string at = " (problem " + std::to_string(index) + ")";
    if (s.ndim != 2 && s.ndim != 3) fail(SFEM_ERR_DIMENSION, "spec: ndim must be 2 or 3" + at);
    for (int d = 0; d < s.ndim; ++d) {
        if (s.dims[d] < 1) fail(SFEM_ERR_ARGUMENT, "spec: at least one element per direction" + at);
        if (!(s.box[2 * d + 1] > s.box[2 * d])) fail(SFEM_ERR_ARGUMENT, "spec: box extents must be increasing" + at);
    }
    if (s.clamp_axis < 0 || s.clamp_axis >= s.ndim || s.load_axis < 0 || s.load_axis >= s.ndim ||
        (s.clamp_side != 0 && s.clamp_side != 1) || (s.load_side != 0 && s.load_side != 1))
        fail(SFEM_ERR_ARGUMENT, "spec: clamp / load face must be an axis < ndim and a side 0 or 1" + at);
    if (!(s.young_lo > 0.0) || !(s.young_hi >= s.young_lo) || !std::isfinite(s.young_hi))
        fail(SFEM_ERR_ARGUMENT, "spec: need 0 < young_lo <= young_hi" + at);
    if (!(s.nu > -1.0 && s.nu < 0.5)) fail(SFEM_ERR_ARGUMENT, "spec: nu must lie in (-1, 1/2)" + at);
}

int64_t spec_nodes(const sfem_problem_spec& s) {
    int64_t n = 1;
    for (int d = 0; d < s.ndim; ++d) n *= s.dims[d] + 1;
    return n;
}

int64_t spec_elems(const sfem_problem_spec& s) {
    int64_t n = 1;
    for (int d = 0; d < s.ndim; ++d) n *= s.dims[d];
    return n;
}

}  // namespace

int64_t spec_free_dofs(const sfem_problem_spec& s) {
    int64_t face = 1;
    for (int d = 0; d < s.ndim; ++d)
        if (d != s.clamp_axis) face *= s.dims[d] + 1;
    return (spec_nodes(s) - face) * s.ndim;
}

// The problems concatenated (node ids offset per problem) as one device mesh
// and its loads / clamp.  One problem: the mesh is declared its grid.
void spec_build(sfem_ctx* ctx, int64_t nprob, const sfem_problem_spec* specs, sfem_mesh** mesh_out,
                sfem_bc** bc_out, std::vector<int64_t>* node_off) {
    if (nprob < 1 || !specs) fail(SFEM_ERR_ARGUMENT, "spec: no problems");
    const int nd = specs[0].ndim;
    std::vector<SpecDev> hs((size_t)nprob);
    std::vector<int64_t> noff((size_t)nprob + 1, 0), eoff((size_t)nprob + 1, 0);
    for (int64_t p = 0; p < nprob; ++p) {
        const sfem_problem_spec& s = specs[p];
        check_spec(s, p);
        if (s.ndim != nd) fail(SFEM_ERR_ARGUMENT, "spec: the problems of a batch share one dimension");
        SpecDev& D = hs[(size_t)p];
        D = SpecDev{};
        D.ndim = nd;
        D.clamp_axis = s.clamp_axis;
        D.clamp_side = s.clamp_side;
        D.load_axis = s.load_axis;
        D.load_side = s.load_side;
        for (int d = 0; d < 3; ++d) {
            D.ne[d] = d < nd ? s.dims[d] : 1;
            D.nn[d] = d < nd ? s.dims[d] + 1 : 1;
            D.trac[d] = d < nd ? s.traction[d] : 0.0;
            D.body[d] = d < nd ? s.body[d] : 0.0;
        }
        for (int k = 0; k < 6; ++k) D.box[k] = k < 2 * nd ? s.box[k] : 0.0;
        D.ylo = s.young_lo;
        D.yhi = s.young_hi;
        D.nu = s.nu;
        D.noise = s.noise;
        D.seed = s.seed;
        D.node_off = noff[(size_t)p];
        D.elem_off = eoff[(size_t)p];
        noff[(size_t)p + 1] = noff[(size_t)p] + spec_nodes(s);
        eoff[(size_t)p + 1] = eoff[(size_t)p] + spec_elems(s);
    }
    const int64_t N = noff[(size_t)nprob], E = eoff[(size_t)nprob];
    if (N * nd >= INT32_MAX) fail(SFEM_ERR_UNSUPPORTED, "spec: 2^31 or more dofs");
    const int nen = 1 << nd;
    std::unique_ptr<sfem_mesh> m(new sfem_mesh());
    m->ctx = ctx;
    m->ndim = nd;
    m->cls = SFEM_Q1;
    m->ndof = nd;
    m->nen = nen;
    m->n_nodes = N;
    m->n_elem = E;
    m->ien_rows = E;
    m->id_rows = N;
    m->n_young = E;
    m->coords.reset(ctx, (size_t)(nd * N));
    m->dof.reset(ctx, (size_t)(nd * N));
    m->ien.reset(ctx, (size_t)(nen * E));
    m->young.reset(ctx, (size_t)E);
    m->nu_e.reset(ctx, (size_t)E);
    std::unique_ptr<sfem_bc> bc(new sfem_bc());
    bc->ctx = ctx;
    bc->n = N * nd;
    bc->f.reset(ctx, (size_t)(N * nd));
    bc->fixed.reset(ctx, (size_t)(N * nd));
    bc->qd.reset(ctx, (size_t)(N * nd));
    bc->qd_zero = true;
    memset_d(ctx, bc->qd.get(), 0, sizeof(double) * N * nd);
    DevBuf<SpecDev> ds(ctx, (size_t)nprob);
    DevBuf<int64_t> dn(ctx, (size_t)nprob + 1), de(ctx, (size_t)nprob + 1);
    h2d(ctx, ds.get(), hs.data(), sizeof(SpecDev) * nprob);
    h2d(ctx, dn.get(), noff.data(), sizeof(int64_t) * (nprob + 1));
    h2d(ctx, de.get(), eoff.data(), sizeof(int64_t) * (nprob + 1));
    hipLaunchKernelGGL(spec_nodes_kernel, dim3(grid_for(N, 256)), dim3(256), 0, ctx->stream, ds.get(), (int)nprob,
                       dn.get(), N, m->coords.get(), m->dof.get(), bc->f.get(), bc->fixed.get());
    SFEM_LAUNCH_CHECK();
    hipLaunchKernelGGL(spec_elems_kernel, dim3(grid_for(E, 256)), dim3(256), 0, ctx->stream, ds.get(), (int)nprob,
                       de.get(), E, m->ien.get(), m->young.get(), m->nu_e.get());
    SFEM_LAUNCH_CHECK();
    if (nprob == 1) {  // the generated mesh is its grid (meshgrid numbering and corner order)
        m->grid_nd = nd;
        for (int d = 0; d < 3; ++d) m->grid_ne[d] = d < nd ? specs[0].dims[d] : 1;
        for (int a = 0; a < nen; ++a) {
            const int code = (q1_sx(a) > 0) + 2 * (q1_sy(a) > 0) + 4 * (q1_sz(a) > 0 && nd == 3);
            m->grid_corner[code] = a;
        }
    }
    sync(ctx);  // the staging buffers are released below
    if (node_off) *node_off = noff;
    *mesh_out = m.release();
    *bc_out = bc.release();
}

}  // namespace sfem

using namespace sfem;

extern "C" {

int sfem_spec_build(sfem_ctx* ctx, const sfem_problem_spec* spec, sfem_mesh** mesh, sfem_bc** bc) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!spec || !mesh || !bc) { ctx->last_error = "spec_build: NULL argument"; return SFEM_ERR_ARGUMENT; }
    *mesh = nullptr;
    *bc = nullptr;
    sfem_mesh* m = nullptr;
    sfem_bc* b = nullptr;
    const int st = guard(ctx, [&] { spec_build(ctx, 1, spec, &m, &b, nullptr); });
    if (st) return st;
    ctx_retain(ctx);
    ctx_retain(ctx);
    *mesh = m;
    *bc = b;
    return SFEM_OK;
}

int sfem_spec_export(sfem_ctx* ctx, const sfem_problem_spec* spec, double* node_list, double* young, double* f,
                     uint8_t* fixed) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!spec) { ctx->last_error = "spec_export: NULL spec"; return SFEM_ERR_ARGUMENT; }
    return guard(ctx, [&] {
        sfem_mesh* m = nullptr;
        sfem_bc* b = nullptr;
        spec_build(ctx, 1, spec, &m, &b, nullptr);
        std::unique_ptr<sfem_mesh> mh(m);
        std::unique_ptr<sfem_bc> bh(b);
        if (node_list) d2h(ctx, node_list, m->coords.get(), sizeof(double) * m->ndim * m->n_nodes);
        if (young) d2h(ctx, young, m->young.get(), sizeof(double) * m->n_elem);
        if (f) d2h(ctx, f, b->f.get(), sizeof(double) * b->n);
        if (fixed) d2h(ctx, fixed, b->fixed.get(), (size_t)b->n);
    });
}

int sfem_batch_create_specs(sfem_ctx* ctx, int64_t nprob, const sfem_problem_spec* specs, sfem_batch** out) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!out || !specs) { ctx->last_error = "batch_create_specs: NULL argument"; return SFEM_ERR_ARGUMENT; }
    *out = nullptr;
    sfem_batch* B = nullptr;
    const int st = guard(ctx, [&] {
        sfem_mesh* m = nullptr;
        sfem_bc* b = nullptr;
        std::vector<int64_t> noff;
        spec_build(ctx, nprob, specs, &m, &b, &noff);
        ctx_retain(ctx);  // the mesh and bc are released through their destroy calls
        ctx_retain(ctx);
        std::vector<int64_t> nfree((size_t)nprob);
        for (int64_t p = 0; p < nprob; ++p) nfree[(size_t)p] = spec_free_dofs(specs[p]);
        B = batch_from_device(ctx, nprob, m, b, noff, nfree);
    });
    if (st) return st;
    ctx_retain(ctx);
    *out = B;
    return SFEM_OK;
}

}  // extern "C"
