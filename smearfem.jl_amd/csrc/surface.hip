// surface.hip -- the boundary machinery of the reference example on the device:
//   * the surface penalty b of apply_boundary_conditions
//     (examples/vector3D.jl:175-264) as a face element of the assembly
//     pipeline (same keys, stable sort and in-order segment sums as K),
//   * K_bar = K + beta*b (examples/vector3D.jl:308),
//   * structured meshes (meshgrid, examples/vector3D.jl:10-130) with the
//     cylinder map of inflate_sphere (src/PostProcess.jl:30-46),
//   * Dirichlet constraints from coordinate planes (setboundaryCond,
//     examples/vector3D.jl:133-173),
//   * new loads / prescribed values for an eliminated system (the sweep,
//     examples/vector3D.jl:310-322; kernel in eliminate.hip).
#include <algorithm>
#include <cstring>
#include <vector>

#include "basis.hpp"
#include "sfem_objects.hpp"

// The penalty values are compared bitwise with a plain C restatement: every
// product is rounded on its own.
#pragma clang fp contract(off)

namespace sfem {

namespace {

// ------------------------------------------------------- surface penalty --
// One thread per (face, a, b): the 3x3 block of b for face nodes a, b summed
// over the four Gauss points in the reference's order (:203-208).  Per point:
// dxdξ = coords * ΔN (:219-220), w = |dxdξ[:,1] x dxdξ[:,2]| (:222-223, the
// weights are 1), be = N_a N_b on equal dof components and 0 elsewhere
// (M'M, :225-231).  V layout as the volume kernels:
// V[face*144 + (a*4 + b)*9 + i*3 + j].
__global__ __launch_bounds__(256) void face_mass_kernel(const double* __restrict__ coords,
                                                        const int32_t* __restrict__ ien, int64_t n_faces,
                                                        double* __restrict__ V) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_faces * 16) return;
    const int64_t e = t >> 4;
    const int pair = (int)(t & 15), a = pair >> 2, b = pair & 3;
    double c[4][3];
    for (int q = 0; q < 4; ++q) {
        const int32_t node = ien[e * 4 + q];
        if (node < 0) return;  // out of NodeList: reported by the validation pass
        for (int d = 0; d < 3; ++d) c[q][d] = coords[(int64_t)node * 3 + d];
    }
    double acc = 0.0;
    for (int g = 0; g < 4; ++g) {
        double xi, eta, zeta;
        gauss_point(g, xi, eta, zeta);
        double t1[3] = {0.0, 0.0, 0.0}, t2[3] = {0.0, 0.0, 0.0};
        for (int q = 0; q < 4; ++q) {
            double dn[2];
            q1_2d_dN(q, xi, eta, dn);
            for (int d = 0; d < 3; ++d) {
                t1[d] += c[q][d] * dn[0];
                t2[d] += c[q][d] * dn[1];
            }
        }
        const double x = t1[1] * t2[2] - t1[2] * t2[1];
        const double y = t1[2] * t2[0] - t1[0] * t2[2];
        const double z = t1[0] * t2[1] - t1[1] * t2[0];
        const double w = 1.0 * sqrt(x * x + y * y + z * z);
        acc += w * (q1_2d_N(a, xi, eta) * q1_2d_N(b, xi, eta));
    }
    double* out = V + e * 144 + pair * 9;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) out[i * 3 + j] = i == j ? acc : 0.0;
}

// Julia column-major n x k (1-based) -> element-major int32 (0-based, -1 out of
// range), rows [row0, row0 + n) of dst.
__global__ void face_ien_kernel(const int64_t* __restrict__ src, int64_t n, int k, int64_t n_nodes, int64_t row0,
                                int32_t* __restrict__ dst) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * k) return;
    const int64_t e = t / k;
    const int a = (int)(t - e * k);
    const int64_t v = src[(int64_t)a * n + e];
    dst[(row0 + e) * k + a] = (v >= 1 && v <= n_nodes) ? (int32_t)(v - 1) : -1;
}

__global__ void id_kernel(const int64_t* __restrict__ src, int64_t rows, int ndof, int32_t* __restrict__ dst) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= rows * ndof) return;
    const int64_t node = t / ndof;
    const int d = (int)(t - node * ndof);
    const int64_t v = src[(int64_t)d * rows + node];
    dst[t] = v < 1 ? -1 : (v > INT32_MAX ? -2 : (int32_t)(v - 1));
}

// ------------------------------------------------------------ K + beta b --
// Block CSC columns merged by row unit (both sorted ascending).
__global__ void add_count_kernel(const int64_t* __restrict__ ap, const int32_t* __restrict__ ar,
                                 const int64_t* __restrict__ bp, const int32_t* __restrict__ br, int64_t nunits,
                                 int64_t* __restrict__ cnt) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nunits) return;
    int64_t i = ap[c], j = bp[c], n = 0;
    const int64_t ie = ap[c + 1], je = bp[c + 1];
    while (i < ie || j < je) {
        if (j >= je || (i < ie && ar[i] < br[j])) ++i;
        else if (i >= ie || br[j] < ar[i]) ++j;
        else { ++i; ++j; }
        ++n;
    }
    cnt[c] = n;
}

__global__ void add_fill_kernel(const int64_t* __restrict__ ap, const int32_t* __restrict__ ar,
                                const double* __restrict__ av, const int64_t* __restrict__ bp,
                                const int32_t* __restrict__ br, const double* __restrict__ bv, double beta, int bs2,
                                int64_t nunits, const int64_t* __restrict__ cp, int32_t* __restrict__ cr,
                                int32_t* __restrict__ cc, double* __restrict__ cv) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nunits) return;
    int64_t i = ap[c], j = bp[c], o = cp[c];
    const int64_t ie = ap[c + 1], je = bp[c + 1];
    while (i < ie || j < je) {
        if (j >= je || (i < ie && ar[i] < br[j])) {
            cr[o] = ar[i];
            for (int v = 0; v < bs2; ++v) cv[o * bs2 + v] = av[i * bs2 + v] + 0.0;
            ++i;
        } else if (i >= ie || br[j] < ar[i]) {
            cr[o] = br[j];
            for (int v = 0; v < bs2; ++v) cv[o * bs2 + v] = 0.0 + beta * bv[j * bs2 + v];
            ++j;
        } else {
            cr[o] = ar[i];
            for (int v = 0; v < bs2; ++v) cv[o * bs2 + v] = av[i * bs2 + v] + beta * bv[j * bs2 + v];
            ++i;
            ++j;
        }
        cc[o] = (int32_t)c;
        ++o;
    }
}

// -------------------------------------------------------------- meshgrid --
// numpy / Julia-style linspace point i of ne intervals, end point exact.
__device__ __forceinline__ double lin(double x0, double x1, int64_t ne, int64_t i) {
    if (i == ne) return x1;
    const double step = (x1 - x0) / (double)ne;
    return (double)i * step + x0;
}

__global__ void grid_nodes_kernel(int ndim, int64_t nx, int64_t ny, int64_t nz, double x0, double x1, double y0,
                                  double y1, double z0, double z1, int inflate, int64_t n_nodes,
                                  double* __restrict__ coords, int32_t* __restrict__ dof) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n_nodes) return;
    const int64_t i = m % (nx + 1), j = (m / (nx + 1)) % (ny + 1), k = m / ((nx + 1) * (ny + 1));
    double x = lin(x0, x1, nx, i), y = lin(y0, y1, ny, j);
    if (inflate) {
        const double cx = 0.5 * (x0 + x1), cy = 0.5 * (y0 + y1);
        const double dx = x - cx, dy = y - cy;
        const double scale = fmax(fabs(dx), fabs(dy));
        if (scale == 0.0) {
            x = 0.0;
            y = 0.0;
        } else {
            const double r = sqrt(dx * dx + dy * dy);
            x = scale * dx / r;
            y = scale * dy / r;
        }
    }
    coords[m * ndim + 0] = x;
    coords[m * ndim + 1] = y;
    if (ndim == 3) coords[m * ndim + 2] = lin(z0, z1, nz, k);
    for (int l = 0; l < ndim; ++l) dof[m * ndim + l] = (int32_t)(ndim * m + l);
}

// IEN in the reference's element order and node order (:40-56, :88-120);
// faces: bottom (z = z0) then top (z = z1) lists, both in (j, i) order.  With
// a single element layer the reference fills only the bottom list (if/elseif);
// the top list stays 0, which is out of range: -1 here.
__global__ void grid_elems_kernel(int ndim, int64_t nx, int64_t ny, int64_t nz, int64_t n_elem,
                                  int32_t* __restrict__ ien, int32_t* __restrict__ faces) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_elem) return;
    const int64_t row = nx + 1;
    if (ndim == 2) {
        const int64_t i = n % nx, j = n / nx;
        const int64_t n0 = j * row + i;
        int32_t* e = ien + n * 4;
        e[0] = (int32_t)n0;
        e[1] = (int32_t)(n0 + 1);
        e[2] = (int32_t)(n0 + row + 1);
        e[3] = (int32_t)(n0 + row);
        return;
    }
    const int64_t i = n % nx, j = (n / nx) % ny, k = n / (nx * ny);
    const int64_t p = (nx + 1) * (ny + 1);
    const int64_t n0 = k * p + j * row + i;
    int32_t* e = ien + n * 8;
    e[0] = (int32_t)n0;
    e[1] = (int32_t)(n0 + 1);
    e[2] = (int32_t)(n0 + row + 1);
    e[3] = (int32_t)(n0 + row);
    for (int a = 0; a < 4; ++a) e[4 + a] = (int32_t)(e[a] + p);
    if (!faces) return;
    const int64_t nf = nx * ny, f = j * nx + i;
    if (k == 0)
        for (int a = 0; a < 4; ++a) faces[f * 4 + a] = e[a];
    if (k == nz - 1) {
        for (int a = 0; a < 4; ++a) faces[(nf + f) * 4 + a] = k == 0 ? -1 : e[4 + a];
    }
}

// ------------------------------------------------------- plane constraints --
__global__ void bc_planes_kernel(const double* __restrict__ coords, int ndim, const int32_t* __restrict__ dof,
                                 int ndof, int64_t n_nodes, int nrules, const int* __restrict__ axis,
                                 const double* __restrict__ coord, const int* __restrict__ rdof,
                                 const double* __restrict__ value, int64_t n_total, uint8_t* __restrict__ fixed,
                                 double* __restrict__ qd, int* __restrict__ bad) {
    const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= n_nodes) return;
    unsigned done = 0;
    for (int r = 0; r < nrules; ++r) {
        const int d = rdof[r];
        if ((done >> d) & 1u) continue;
        if (coords[node * ndim + axis[r]] == coord[r]) {
            done |= 1u << d;
            const int32_t g = dof[node * ndof + d];
            if (g < 0 || g >= n_total) {
                atomicOr(bad, 1);
                continue;
            }
            fixed[g] = 1;
            qd[g] = value[r];
        }
    }
}

// y = K x walking K's block columns as rows: exact for the assembled K and b
// (bitwise symmetric) and for an uploaded CSC (which stores K').  One thread
// per unit; the sum runs over the column's blocks in ascending row order.
__global__ void csr_apply_kernel(const int64_t* __restrict__ bcolptr, const int32_t* __restrict__ browidx,
                                 const double* __restrict__ bval, int64_t nunits, int bs, const double* __restrict__ x,
                                 double* __restrict__ y) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nunits) return;
    const int bs2 = bs * bs;
    for (int j = 0; j < bs; ++j) {
        double acc = 0.0;
        for (int64_t e = bcolptr[u]; e < bcolptr[u + 1]; ++e) {
            const int64_t r = browidx[e];
            for (int i = 0; i < bs; ++i) acc += bval[e * bs2 + i * bs + j] * x[r * bs + i];
        }
        y[u * bs + j] = acc;
    }
}

template <typename T>
T* new_or_null() {
    return new (std::nothrow) T();
}

// Surface mesh from device coordinates / ids: elements = bottom then top.
void surface_from_device(sfem_ctx* ctx, sfem_mesh* f, const double* coords, const int32_t* dof, int64_t n_nodes,
                         int64_t id_rows, int64_t n_faces) {
    f->ctx = ctx;
    f->ndim = 3;
    f->cls = SFEM_Q1;
    f->ndof = 3;
    f->nen = 4;
    f->n_nodes = n_nodes;
    f->n_elem = 2 * n_faces;
    f->ien_rows = 2 * n_faces;
    f->id_rows = id_rows;
    f->n_young = 1;
    f->face = true;
    f->key_period = n_faces;
    f->coords.reset(ctx, (size_t)(3 * n_nodes));
    SFEM_HIP(hipMemcpyAsync(f->coords.get(), coords, sizeof(double) * 3 * n_nodes, hipMemcpyDeviceToDevice,
                            ctx->stream));
    f->dof.reset(ctx, (size_t)(3 * id_rows));
    SFEM_HIP(hipMemcpyAsync(f->dof.get(), dof, sizeof(int32_t) * 3 * id_rows, hipMemcpyDeviceToDevice, ctx->stream));
    const double one = 1.0;
    f->young.reset(ctx, 1);
    h2d(ctx, f->young.get(), &one, sizeof(double));
}

}  // namespace

void launch_face_mass(sfem_ctx* ctx, const sfem_mesh* m, double* V) {
    if (m->n_elem == 0) return;
    hipLaunchKernelGGL(face_mass_kernel, dim3(grid_for(m->n_elem * 16, 256)), dim3(256), 0, ctx->stream,
                       m->coords.get(), m->ien.get(), m->n_elem, V);
    SFEM_LAUNCH_CHECK();
}

}  // namespace sfem

using namespace sfem;

extern "C" {

int sfem_surface_mesh_create(sfem_ctx* ctx, int ndim, int n_dof, const double* node_list, int64_t n_nodes,
                             const int64_t* ien_btm, const int64_t* ien_top, int64_t n_faces, int nfn,
                             const int64_t* id, int64_t id_rows, sfem_mesh** out) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!out) { ctx->last_error = "out is NULL"; return SFEM_ERR_ARGUMENT; }
    *out = nullptr;
    sfem_mesh* m = new_or_null<sfem_mesh>();
    if (!m) return SFEM_ERR_MEMORY;
    int st = guard(ctx, [&] {
        if (n_faces < 0 || n_nodes < 0 || id_rows < 0 || nfn <= 0 ||
            (n_faces > 0 && (!ien_btm || !ien_top || !node_list || !id)))
            fail(SFEM_ERR_ARGUMENT, "surface_mesh_create: bad sizes or NULL arrays");
        if (n_nodes >= INT32_MAX || id_rows >= INT32_MAX) fail(SFEM_ERR_UNSUPPORTED, "2^31 or more nodes");
        if (n_faces > 0 && (ndim != 3 || nfn != 4 || n_dof != 3)) {
            // the reference gathers face 1's coordinates first (:199-200)
            for (int a = 0; a < nfn; ++a) {
                const int64_t vt = ien_top[(int64_t)a * n_faces], vb = ien_btm[(int64_t)a * n_faces];
                if (vt < 1 || vt > n_nodes || vb < 1 || vb > n_nodes)
                    fail(SFEM_ERR_BOUNDS, "apply_boundary_conditions: BoundsError at face 1");
            }
            if (ndim != 2 && ndim != 3) fail(SFEM_ERR_UNDEFVAR, "apply_boundary_conditions: UndefVarError at face 1");
            fail(SFEM_ERR_DIMENSION, "apply_boundary_conditions: DimensionMismatch at face 1 (3D Q4 faces, "
                                     "nDof = 3 only)");
        }
        m->ctx = ctx;
        m->ndim = 3;
        m->cls = SFEM_Q1;
        m->ndof = 3;
        m->nen = 4;
        m->n_nodes = n_nodes;
        m->n_elem = 2 * n_faces;
        m->ien_rows = 2 * n_faces;
        m->id_rows = id_rows;
        m->n_young = 1;
        m->face = true;
        m->key_period = n_faces;
        if (n_faces == 0) return;
        m->coords.reset(ctx, (size_t)(3 * n_nodes));
        h2d(ctx, m->coords.get(), node_list, sizeof(double) * 3 * n_nodes);
        m->ien.reset(ctx, (size_t)(2 * n_faces * 4));
        {
            DevBuf<int64_t> raw(ctx, (size_t)(n_faces * 4));
            const dim3 g(grid_for(n_faces * 4, 256));
            h2d(ctx, raw.get(), ien_btm, sizeof(int64_t) * n_faces * 4);
            hipLaunchKernelGGL(face_ien_kernel, g, dim3(256), 0, ctx->stream, raw.get(), n_faces, 4, n_nodes,
                               (int64_t)0, m->ien.get());
            SFEM_LAUNCH_CHECK();
            sync(ctx);
            h2d(ctx, raw.get(), ien_top, sizeof(int64_t) * n_faces * 4);
            hipLaunchKernelGGL(face_ien_kernel, g, dim3(256), 0, ctx->stream, raw.get(), n_faces, 4, n_nodes,
                               n_faces, m->ien.get());
            SFEM_LAUNCH_CHECK();
            sync(ctx);
        }
        if (id_rows > 0) {
            DevBuf<int64_t> raw(ctx, (size_t)(id_rows * 3));
            h2d(ctx, raw.get(), id, sizeof(int64_t) * id_rows * 3);
            m->dof.reset(ctx, (size_t)(id_rows * 3));
            hipLaunchKernelGGL(id_kernel, dim3(grid_for(id_rows * 3, 256)), dim3(256), 0, ctx->stream, raw.get(),
                               id_rows, 3, m->dof.get());
            SFEM_LAUNCH_CHECK();
            sync(ctx);
        }
        const double one = 1.0;
        m->young.reset(ctx, 1);
        h2d(ctx, m->young.get(), &one, sizeof(double));
        sync(ctx);
    });
    if (st) { delete m; return st; }
    ctx_retain(ctx);
    *out = m;
    return SFEM_OK;
}

int sfem_assemble_surface(sfem_ctx* ctx, int ndim, int n_dof, const double* node_list, int64_t n_nodes,
                          const int64_t* ien_btm, const int64_t* ien_top, int64_t n_faces, int nfn,
                          const int64_t* id, int64_t id_rows, sfem_csr** out) {
    sfem_mesh* m = nullptr;
    int st = sfem_surface_mesh_create(ctx, ndim, n_dof, node_list, n_nodes, ien_btm, ien_top, n_faces, nfn, id,
                                      id_rows, &m);
    if (st) return st;
    st = sfem_assemble_mesh(ctx, m, SFEM_ASM_DEFAULT, out);
    sfem_mesh_destroy(m);
    return st;
}

int sfem_csr_add(sfem_ctx* ctx, const sfem_csr* K, double beta, const sfem_csr* b, sfem_csr** out) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!K || !b || !out) { ctx->last_error = "csr_add: NULL argument"; return SFEM_ERR_ARGUMENT; }
    *out = nullptr;
    sfem_csr* C = new_or_null<sfem_csr>();
    if (!C) return SFEM_ERR_MEMORY;
    int st = guard(ctx, [&] {
        if (K->m != b->m || K->n != b->n)
            fail(SFEM_ERR_DIMENSION, "csr_add: DimensionMismatch (matrix sizes differ)");
        if (K->bs != b->bs || K->transposed != b->transposed || K->nunits != b->nunits)
            fail(SFEM_ERR_UNSUPPORTED, "csr_add: both matrices need the same block layout");
        const int bs = K->bs, bs2 = bs * bs;
        const int64_t nu = K->nunits;
        C->ctx = ctx;
        C->m = K->m;
        C->n = K->n;
        C->bs = bs;
        C->transposed = K->transposed;
        C->nunits = nu;
        C->bcolptr.reset(ctx, (size_t)nu + 1);
        if (nu == 0) {
            memset_d(ctx, C->bcolptr.get(), 0, sizeof(int64_t));
            sync(ctx);
            return;
        }
        DevBuf<int64_t> cnt(ctx, (size_t)nu + 1);
        memset_d(ctx, cnt.get() + nu, 0, sizeof(int64_t));
        const dim3 g(grid_for(nu, 256));
        hipLaunchKernelGGL(add_count_kernel, g, dim3(256), 0, ctx->stream, K->bcolptr.get(), K->browidx.get(),
                           b->bcolptr.get(), b->browidx.get(), nu, cnt.get());
        SFEM_LAUNCH_CHECK();
        const int64_t nnzb = exclusive_scan_i64(ctx, cnt.get(), C->bcolptr.get(), nu + 1);
        if (nnzb >= INT32_MAX) fail(SFEM_ERR_UNSUPPORTED, "csr_add: 2^31 or more blocks");
        C->nnzb = nnzb;
        C->nnz = nnzb * bs2;
        C->browidx.reset(ctx, (size_t)nnzb);
        C->bcolidx.reset(ctx, (size_t)nnzb);
        C->bval.reset(ctx, (size_t)nnzb * bs2);
        hipLaunchKernelGGL(add_fill_kernel, g, dim3(256), 0, ctx->stream, K->bcolptr.get(), K->browidx.get(),
                           K->bval.get(), b->bcolptr.get(), b->browidx.get(), b->bval.get(), beta, bs2, nu,
                           C->bcolptr.get(), C->browidx.get(), C->bcolidx.get(), C->bval.get());
        SFEM_LAUNCH_CHECK();
        sync(ctx);
    });
    if (st) { delete C; return st; }
    ctx_retain(ctx);
    *out = C;
    return SFEM_OK;
}

int sfem_meshgrid(sfem_ctx* ctx, int ndim, const int64_t* ne, const double* box, int flags, const double* young,
                  int64_t n_young, double nu, sfem_mesh** mesh, sfem_mesh** faces) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!mesh || !ne || !box || !young) { ctx->last_error = "meshgrid: NULL argument"; return SFEM_ERR_ARGUMENT; }
    *mesh = nullptr;
    if (faces) *faces = nullptr;
    sfem_mesh* m = new_or_null<sfem_mesh>();
    sfem_mesh* fm = faces ? new_or_null<sfem_mesh>() : nullptr;
    if (!m || (faces && !fm)) { delete m; delete fm; return SFEM_ERR_MEMORY; }
    int st = guard(ctx, [&] {
        if (ndim != 2 && ndim != 3) fail(SFEM_ERR_DIMENSION, "meshgrid: ndim must be 2 or 3");
        if (faces && ndim != 3) fail(SFEM_ERR_DIMENSION, "meshgrid: face lists are built for 3D meshes");
        const int64_t nx = ne[0], ny = ne[1], nz = ndim == 3 ? ne[2] : 1;
        if (nx < 1 || ny < 1 || nz < 1) fail(SFEM_ERR_ARGUMENT, "meshgrid: at least one element per direction");
        const int64_t n_nodes = (nx + 1) * (ny + 1) * (ndim == 3 ? nz + 1 : 1);
        const int64_t n_elem = nx * ny * (ndim == 3 ? nz : 1);
        if (n_nodes * ndim >= INT32_MAX) fail(SFEM_ERR_UNSUPPORTED, "meshgrid: 2^31 or more dofs");
        if (n_young != 1 && n_young != n_elem) fail(SFEM_ERR_ARGUMENT, "Young: need 1 or n_elem values");
        const int nen = 1 << ndim;
        m->ctx = ctx;
        m->ndim = ndim;
        m->cls = SFEM_Q1;
        m->ndof = ndim;
        m->nen = nen;
        m->n_nodes = n_nodes;
        m->n_elem = n_elem;
        m->ien_rows = n_elem;
        m->id_rows = n_nodes;
        m->n_young = n_young;
        m->nu = nu;
        m->coords.reset(ctx, (size_t)(ndim * n_nodes));
        m->dof.reset(ctx, (size_t)(ndim * n_nodes));
        m->ien.reset(ctx, (size_t)(nen * n_elem));
        m->young.reset(ctx, (size_t)n_young);
        h2d(ctx, m->young.get(), young, sizeof(double) * n_young);
        const double z0 = ndim == 3 ? box[4] : 0.0, z1 = ndim == 3 ? box[5] : 0.0;
        hipLaunchKernelGGL(grid_nodes_kernel, dim3(grid_for(n_nodes, 256)), dim3(256), 0, ctx->stream, ndim, nx, ny,
                           nz, box[0], box[1], box[2], box[3], z0, z1, (flags & SFEM_MESH_INFLATE) ? 1 : 0, n_nodes,
                           m->coords.get(), m->dof.get());
        SFEM_LAUNCH_CHECK();
        DevBuf<int32_t> fien;
        const int64_t nf = nx * ny;
        if (faces) fien.reset(ctx, (size_t)(2 * nf * 4));
        hipLaunchKernelGGL(grid_elems_kernel, dim3(grid_for(n_elem, 256)), dim3(256), 0, ctx->stream, ndim, nx, ny,
                           nz, n_elem, m->ien.get(), faces ? fien.get() : nullptr);
        SFEM_LAUNCH_CHECK();
        if (faces) {
            surface_from_device(ctx, fm, m->coords.get(), m->dof.get(), n_nodes, n_nodes, nf);
            fm->ien = std::move(fien);
        }
        const int64_t dims[3] = {nx, ny, nz};
        mesh_set_grid(ctx, m, ndim, dims);  // the generated mesh is its grid
        sync(ctx);
    });
    if (st) { delete m; delete fm; return st; }
    ctx_retain(ctx);
    *mesh = m;
    if (faces) {
        ctx_retain(ctx);
        *faces = fm;
    }
    return SFEM_OK;
}

int sfem_mesh_info(const sfem_mesh* m, int* ndim, int* nen, int* n_dof, int64_t* n_nodes, int64_t* n_elem) {
    if (!m) return SFEM_ERR_ARGUMENT;
    if (ndim) *ndim = m->ndim;
    if (nen) *nen = m->nen;
    if (n_dof) *n_dof = m->ndof;
    if (n_nodes) *n_nodes = m->n_nodes;
    if (n_elem) *n_elem = m->n_elem;
    return SFEM_OK;
}

int sfem_mesh_export(sfem_ctx* ctx, const sfem_mesh* m, double* node_list, int64_t* ien, int64_t* id) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!m) { ctx->last_error = "mesh_export: NULL mesh"; return SFEM_ERR_ARGUMENT; }
    return guard(ctx, [&] {
        if (node_list && m->n_nodes)
            d2h(ctx, node_list, m->coords.get(), sizeof(double) * m->ndim * m->n_nodes);
        if (ien && m->n_elem) {
            std::vector<int32_t> h((size_t)(m->n_elem * m->nen));
            d2h(ctx, h.data(), m->ien.get(), sizeof(int32_t) * h.size());
            for (int64_t e = 0; e < m->n_elem; ++e)
                for (int a = 0; a < m->nen; ++a) ien[(int64_t)a * m->n_elem + e] = (int64_t)h[e * m->nen + a] + 1;
        }
        if (id && m->id_rows && m->ndof >= 2) {
            std::vector<int32_t> h((size_t)(m->id_rows * m->ndof));
            d2h(ctx, h.data(), m->dof.get(), sizeof(int32_t) * h.size());
            for (int64_t nd = 0; nd < m->id_rows; ++nd)
                for (int d = 0; d < m->ndof; ++d) id[(int64_t)d * m->id_rows + nd] = (int64_t)h[nd * m->ndof + d] + 1;
        }
    });
}

int sfem_bc_planes(sfem_ctx* ctx, const sfem_mesh* m, int64_t n_total, int nrules, const int* axis,
                   const double* coord, const int* dof, const double* value, const double* f, sfem_bc** out) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!m || !out || (nrules > 0 && (!axis || !coord || !dof || !value))) {
        ctx->last_error = "bc_planes: NULL argument";
        return SFEM_ERR_ARGUMENT;
    }
    *out = nullptr;
    sfem_bc* bc = new_or_null<sfem_bc>();
    if (!bc) return SFEM_ERR_MEMORY;
    int st = guard(ctx, [&] {
        if (n_total < 0 || nrules < 0 || nrules > 64) fail(SFEM_ERR_ARGUMENT, "bc_planes: 0..64 rules");
        if (m->face) fail(SFEM_ERR_ARGUMENT, "bc_planes: needs a volume mesh");
        if (m->ndof < 2 || m->dof.size() == 0) fail(SFEM_ERR_ARGUMENT, "bc_planes: the mesh has no ID (nDof >= 2)");
        std::vector<int> ax(nrules), dd(nrules);
        for (int r = 0; r < nrules; ++r) {
            if (axis[r] < 1 || axis[r] > m->ndim) fail(SFEM_ERR_ARGUMENT, "bc_planes: axis out of 1..ndim");
            if (dof[r] < 1 || dof[r] > m->ndof) fail(SFEM_ERR_ARGUMENT, "bc_planes: dof out of 1..nDof");
            ax[r] = axis[r] - 1;
            dd[r] = dof[r] - 1;
        }
        bc->ctx = ctx;
        bc->n = n_total;
        bc->f.reset(ctx, (size_t)n_total);
        bc->fixed.reset(ctx, (size_t)n_total);
        bc->qd.reset(ctx, (size_t)n_total);
        if (f) h2d(ctx, bc->f.get(), f, sizeof(double) * n_total);
        else memset_d(ctx, bc->f.get(), 0, sizeof(double) * n_total);
        memset_d(ctx, bc->fixed.get(), 0, (size_t)n_total);
        memset_d(ctx, bc->qd.get(), 0, sizeof(double) * n_total);
        if (nrules > 0 && m->n_nodes > 0) {
            DevBuf<int> dax(ctx, (size_t)nrules), ddof(ctx, (size_t)nrules), bad(ctx, 1);
            DevBuf<double> dco(ctx, (size_t)nrules), dval(ctx, (size_t)nrules);
            h2d(ctx, dax.get(), ax.data(), sizeof(int) * nrules);
            h2d(ctx, ddof.get(), dd.data(), sizeof(int) * nrules);
            h2d(ctx, dco.get(), coord, sizeof(double) * nrules);
            h2d(ctx, dval.get(), value, sizeof(double) * nrules);
            memset_d(ctx, bad.get(), 0, sizeof(int));
            const int64_t nn = std::min(m->n_nodes, m->id_rows);
            hipLaunchKernelGGL(bc_planes_kernel, dim3(grid_for(nn, 256)), dim3(256), 0, ctx->stream,
                               m->coords.get(), m->ndim, m->dof.get(), m->ndof, nn, nrules, dax.get(), dco.get(),
                               ddof.get(), dval.get(), n_total, bc->fixed.get(), bc->qd.get(), bad.get());
            SFEM_LAUNCH_CHECK();
            int b = 0;
            d2h(ctx, &b, bad.get(), sizeof(int));
            if (b) fail(SFEM_ERR_BOUNDS, "bc_planes: a constrained dof id is outside 1..n_total");
        }
        sync(ctx);
    });
    if (st) { delete bc; return st; }
    ctx_retain(ctx);
    *out = bc;
    return SFEM_OK;
}

int sfem_csr_apply(sfem_ctx* ctx, const sfem_csr* K, const double* x, double* y) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!K || !x || !y) { ctx->last_error = "csr_apply: NULL argument"; return SFEM_ERR_ARGUMENT; }
    return guard(ctx, [&] {
        if (K->m != K->n || K->m != K->nunits * K->bs) fail(SFEM_ERR_DIMENSION, "csr_apply: square matrices only");
        const int64_t n = K->n;
        if (n == 0) return;
        DevBuf<double> dx(ctx, (size_t)n), dy(ctx, (size_t)n);
        h2d(ctx, dx.get(), x, sizeof(double) * n);
        hipLaunchKernelGGL(csr_apply_kernel, dim3(grid_for(K->nunits, 256)), dim3(256), 0, ctx->stream,
                           K->bcolptr.get(), K->browidx.get(), K->bval.get(), K->nunits, K->bs, dx.get(), dy.get());
        SFEM_LAUNCH_CHECK();
        d2h(ctx, y, dy.get(), sizeof(double) * n);
    });
}

int sfem_system_set_bc(sfem_ctx* ctx, sfem_system* S, const sfem_bc* bc) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!S || !bc) { ctx->last_error = "system_set_bc: NULL argument"; return SFEM_ERR_ARGUMENT; }
    return guard(ctx, [&] { system_set_bc(ctx, S, bc); });
}

}  // extern "C"
