// capi.hip -- extern "C" entry points of include/smearfem.h: context and
// allocator, host scalar utilities, mesh upload and assembly, CSC export.
// Every entry point converts internal exceptions into a status code plus a
// message retrievable with sfem_last_error(); nothing prints or exits.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "basis.hpp"
#include "sfem_objects.hpp"

namespace sfem {

// ------------------------------------------------------------- allocator --
size_t Allocator::size_class(size_t bytes) {
    if (bytes <= 256) return 256;
    // round up to 1/8 steps of the next power of two: bounded waste, few classes
    size_t p = 1;
    while (p < bytes) p <<= 1;
    const size_t step = p >> 3;
    return (bytes + step - 1) / step * step;
}

void* Allocator::alloc(size_t bytes) {
    const size_t sz = size_class(bytes);
    std::lock_guard<std::mutex> lk(mu_);
    auto it = cache_.lower_bound(sz);
    if (it != cache_.end() && it->first == sz) {
        void* p = it->second;
        cache_.erase(it);
        live_[p] = sz;
        in_use_ += (int64_t)sz;
        return p;
    }
    if (limit_ >= 0 && held_ + (int64_t)sz > limit_) {
        for (auto& kv : cache_) { (void)hipFree(kv.second); held_ -= (int64_t)kv.first; }
        cache_.clear();
        if (held_ + (int64_t)sz > limit_) fail(SFEM_ERR_MEMORY, "device memory budget exceeded");
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, sz);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        for (auto& kv : cache_) { (void)hipFree(kv.second); held_ -= (int64_t)kv.first; }
        cache_.clear();
        e = hipMalloc(&p, sz);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            fail(SFEM_ERR_MEMORY, "hipMalloc of " + std::to_string(sz) + " bytes failed");
        }
    }
    live_[p] = sz;
    in_use_ += (int64_t)sz;
    held_ += (int64_t)sz;
    return p;
}

void Allocator::free(void* p) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = live_.find(p);
    if (it == live_.end()) return;
    in_use_ -= (int64_t)it->second;
    cache_.emplace(it->second, p);
    live_.erase(it);
}

void Allocator::release_cached() {
    std::lock_guard<std::mutex> lk(mu_);
    for (auto& kv : cache_) { (void)hipFree(kv.second); held_ -= (int64_t)kv.first; }
    cache_.clear();
}

void Allocator::release_all() {
    release_cached();
    std::lock_guard<std::mutex> lk(mu_);
    for (auto& kv : live_) { (void)hipFree(kv.first); held_ -= (int64_t)kv.second; }
    live_.clear();
    in_use_ = 0;
}

void h2d(sfem_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (bytes) SFEM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
}
void d2h(sfem_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (bytes) SFEM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    SFEM_HIP(hipStreamSynchronize(ctx->stream));
}
void memset_d(sfem_ctx* ctx, void* dst, int v, size_t bytes) {
    if (bytes) SFEM_HIP(hipMemsetAsync(dst, v, bytes, ctx->stream));
}
void sync(sfem_ctx* ctx) { SFEM_HIP(hipStreamSynchronize(ctx->stream)); }

void ctx_retain(sfem_ctx* ctx) {
    if (ctx) ctx->refs.fetch_add(1);
}

void ctx_release(sfem_ctx* ctx) {
    if (!ctx || ctx->refs.fetch_sub(1) != 1) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->bar_ws) ctx->alloc->free(ctx->bar_ws);
    delete ctx->alloc;
    for (hipStream_t& t : ctx->side)
        if (t) {
            (void)hipStreamSynchronize(t);
            (void)hipStreamDestroy(t);
        }
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

hipStream_t side_stream(sfem_ctx* ctx, int i) {
    hipStream_t& t = ctx->side[i % sfem_ctx::kSideStreams];
    if (!t) SFEM_HIP(hipStreamCreateWithFlags(&t, hipStreamNonBlocking));
    return t;
}

namespace {

__global__ void convert_ien_kernel(const int64_t* __restrict__ src, int64_t ld, int nen, int64_t n_elem,
                                   int64_t n_nodes, int32_t* __restrict__ dst) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_elem * nen) return;
    const int64_t e = t / nen;
    const int a = (int)(t - e * nen);
    const int64_t v = src[(int64_t)a * ld + e];
    dst[t] = (v >= 1 && v <= n_nodes) ? (int32_t)(v - 1) : -1;
}

__global__ void convert_id_kernel(const int64_t* __restrict__ src, int64_t rows, int ndof,
                                  int32_t* __restrict__ dst) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= rows * ndof) return;
    const int64_t node = t / ndof;
    const int d = (int)(t - node * ndof);
    const int64_t v = src[(int64_t)d * rows + node];
    dst[t] = v < 1 ? -1 : (v > INT32_MAX ? -2 : (int32_t)(v - 1));
}

// Structural checks the reference meets at the first element (see
// sfem_mesh_create).  Returns (code, step) or code 0.
void structural_check(int ndim, int cls, int ndof, int nen, int* code, int* step) {
    *code = 0;
    *step = 0;
    if (ndim == 1) { *code = SFEM_ERR_DIMENSION; *step = 1; return; }  // coords*ΔN, :192
    if (ndim != 2 && ndim != 3) { *code = SFEM_ERR_UNDEFVAR; *step = 1; return; }  // gp lists undefined
    if (cls != SFEM_Q1 && cls != SFEM_Q2) { *code = SFEM_ERR_UNDEFVAR; *step = 1; return; }
    if (nen != basis_count(cls, ndim)) { *code = SFEM_ERR_DIMENSION; *step = 1; return; }
    if (cls == SFEM_Q2 && ndim == 3) { *code = SFEM_ERR_DIMENSION; *step = 1; return; }
    if (ndof == 1) return;
    if (ndof == 2 && ndim != 2) { *code = SFEM_ERR_DIMENSION; *step = 3; return; }  // B[1,1:nDof:end]=…, :212
    if (ndof == 3 && ndim != 3) { *code = SFEM_ERR_DIMENSION; *step = 3; return; }  // :220
    if (ndof != 2 && ndof != 3) { *code = SFEM_ERR_UNDEFVAR; *step = 3; return; }   // B never assigned
}

// det of the Jacobian of element 0 at the first Gauss point (host check).
double first_gp_det(int ndim, int cls, int nen, const double* node_list, const int64_t* ien, int64_t ld) {
    double xi, eta, zeta;
    gauss_point(0, xi, eta, zeta);
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = 0; a < nen; ++a) {
        const int64_t node = ien[(int64_t)a * ld] - 1;
        double d[3];
        if (ndim == 3) q1_3d_dN(a, xi, eta, zeta, d);
        else if (cls == SFEM_Q2) q2_2d_dN(a, xi, eta, d);
        else q1_2d_dN(a, xi, eta, d);
        for (int i = 0; i < ndim; ++i)
            for (int j = 0; j < ndim; ++j) J[i][j] += node_list[node * ndim + i] * d[j];
    }
    if (ndim == 2) return J[0][0] * J[1][1] - J[0][1] * J[1][0];
    return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) + J[0][1] * (J[1][2] * J[2][0] - J[1][0] * J[2][2]) +
           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}

}  // namespace
}  // namespace sfem

using namespace sfem;

extern "C" {

int sfem_ctx_create(int device, sfem_ctx** out) {
    if (!out) return SFEM_ERR_ARGUMENT;
    *out = nullptr;
    sfem_ctx* ctx = new (std::nothrow) sfem_ctx();
    if (!ctx) return SFEM_ERR_MEMORY;
    int st = guard(nullptr, [&] {
        int ndev = 0;
        SFEM_HIP(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) fail(SFEM_ERR_ARGUMENT, "no such device");
        SFEM_HIP(hipSetDevice(device));
        hipDeviceProp_t prop;
        SFEM_HIP(hipGetDeviceProperties(&prop, device));
        ctx->device = device;
        ctx->n_cu = prop.multiProcessorCount;
        SFEM_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        ctx->alloc = new Allocator(device);
        // launch-shape overrides for measurements (defaults in sfem_ctx)
        const char* nt = std::getenv("SFEM_SPMV_NT");
        if (nt) ctx->tune_nt = nt[0] == '1';
        const char* ws = std::getenv("SFEM_SELL_WAVE_SLICES");
        if (ws) ctx->tune_wave_slices = std::max(0, std::atoi(ws));
        const char* bt = std::getenv("SFEM_BARRIER_TIMEOUT_US");
        if (bt) ctx->bar_timeout_ticks = (unsigned long long)std::max(0LL, std::atoll(bt)) * 100ull;
        const char* bg = std::getenv("SFEM_BARRIER_GROUPED");
        if (bg) ctx->bar_grouped = bg[0] != '0';
        // the no-allocation solve's workspace: each GPU holds its own copy of
        // the array, so its address is taken with this device current
        ctx->scratch = noalloc_scratch_address();
    });
    if (st) { delete ctx; return st; }
    *out = ctx;
    return SFEM_OK;
}

void sfem_ctx_destroy(sfem_ctx* ctx) { ctx_release(ctx); }

const char* sfem_last_error(const sfem_ctx* ctx) { return ctx ? ctx->last_error.c_str() : "no context"; }

int sfem_device_info(sfem_ctx* ctx, char* name, size_t name_len, int* n_cu, int64_t* hbm_bytes) {
    return guard(ctx, [&] {
        hipDeviceProp_t prop;
        SFEM_HIP(hipGetDeviceProperties(&prop, ctx->device));
        if (name && name_len) {
            std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
        }
        if (n_cu) *n_cu = prop.multiProcessorCount;
        if (hbm_bytes) *hbm_bytes = (int64_t)prop.totalGlobalMem;
    });
}

int64_t sfem_ctx_device_bytes(const sfem_ctx* ctx) { return ctx && ctx->alloc ? ctx->alloc->bytes_held() : 0; }

int sfem_ctx_synchronize(sfem_ctx* ctx) {
    return guard(ctx, [&] { sync(ctx); });
}

int sfem_gaussian_quadrature(double a, double b, int n, double* xi, double* w) {
    if (!xi || !w) return SFEM_ERR_ARGUMENT;
    if (n == 2) {
        xi[0] = -(b - a) / (2 * std::sqrt(3.0)) + (b + a) / 2;
        xi[1] = (b - a) / (2 * std::sqrt(3.0)) + (b + a) / 2;
        w[0] = (b - a) / 2;
        w[1] = (b - a) / 2;
        return SFEM_OK;
    }
    if (n == 3) {
        xi[0] = -(b - a) / (2 * std::sqrt(5.0 / 3.0)) + (b + a) / 2;
        xi[1] = 0.0;
        xi[2] = (b - a) / (2 * std::sqrt(5.0 / 3.0)) + (b + a) / 2;
        w[0] = (b - a) / 2 * 5 / 9;
        w[1] = (b - a) / 2 * 8 / 9;
        w[2] = (b - a) / 2 * 5 / 9;
        return SFEM_OK;
    }
    return SFEM_ERR_UNDEFVAR;
}

int sfem_basis_function(int nargs, double xi, double eta, double zeta, int function_class, double* N,
                        double* dN, int* nfun, int* nrow, int* ncol) {
    if (!N || !dN || !nfun || !nrow || !ncol) return SFEM_ERR_ARGUMENT;
    const int nf = basis_count(function_class, nargs);
    if (!nf) return SFEM_ERR_UNDEFVAR;
    *nfun = nf;
    if (function_class == SFEM_Q1 && nargs >= 3) {
        for (int a = 0; a < 8; ++a) {
            double d[3];
            N[a] = q1_3d_N(a, xi, eta, zeta);
            q1_3d_dN(a, xi, eta, zeta, d);
            for (int k = 0; k < 3; ++k) dN[a + 8 * k] = d[k];
        }
        *nrow = 8; *ncol = 3;
    } else if (function_class == SFEM_Q1 && nargs == 2) {
        for (int a = 0; a < 4; ++a) {
            double d[2];
            N[a] = q1_2d_N(a, xi, eta);
            q1_2d_dN(a, xi, eta, d);
            dN[a] = d[0];
            dN[a + 4] = d[1];
        }
        *nrow = 4; *ncol = 2;
    } else if (function_class == SFEM_Q1) {  // 1D: ΔN is the row [-0.5 0.5], :75
        N[0] = 0.5 - 0.5 * xi;
        N[1] = 0.5 + 0.5 * xi;
        dN[0] = -0.5;
        dN[1] = 0.5;
        *nrow = 1; *ncol = 2;
    } else {
        for (int a = 0; a < 9; ++a) {
            double d[2];
            N[a] = q2_2d_N(a, xi, eta);
            q2_2d_dN(a, xi, eta, d);
            dN[a] = d[0];
            dN[a + 9] = d[1];
        }
        *nrow = 9; *ncol = 2;
    }
    return SFEM_OK;
}

int sfem_mesh_create(sfem_ctx* ctx, int ndim, int function_class, int n_dof, const double* node_list,
                     int64_t n_nodes, const int64_t* ien, int64_t ien_rows, int nen, int64_t n_elem,
                     const int64_t* id, int64_t id_rows, const double* young, int64_t n_young, double nu,
                     sfem_mesh** out) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!out) { ctx->last_error = "out is NULL"; return SFEM_ERR_ARGUMENT; }
    *out = nullptr;
    sfem_mesh* m = new (std::nothrow) sfem_mesh();
    if (!m) return SFEM_ERR_MEMORY;
    int st = guard(ctx, [&] {
        if (n_elem < 0 || n_nodes < 0 || ien_rows < 0 || nen <= 0 || (n_elem > 0 && (!ien || !node_list)))
            fail(SFEM_ERR_ARGUMENT, "mesh_create: bad sizes or NULL arrays");
        if (!young || (n_young != 1 && n_young != n_elem)) fail(SFEM_ERR_ARGUMENT, "Young: need 1 or n_elem values");
        if (n_dof >= 2 && !id) fail(SFEM_ERR_ARGUMENT, "ID is required when nDof >= 2");
        if (n_nodes >= INT32_MAX || id_rows >= INT32_MAX) fail(SFEM_ERR_UNSUPPORTED, "2^31 or more nodes");
        m->ctx = ctx;
        m->ndim = ndim; m->cls = function_class; m->ndof = n_dof; m->nen = nen;
        m->n_nodes = n_nodes; m->ien_rows = ien_rows; m->id_rows = n_dof >= 2 ? id_rows : 0;
        m->n_young = n_young; m->nu = nu;
        const int64_t n_used = std::min(n_elem, ien_rows);
        m->n_elem = n_used;
        if (n_elem > ien_rows) { m->pre_code = SFEM_ERR_BOUNDS; m->pre_key = ien_rows * 8 + 0; }
        if (n_elem == 0) return;
        // errors the reference meets at element 1 before per-element checks
        int scode, sstep;
        structural_check(ndim, function_class, n_dof, nen, &scode, &sstep);
        if (scode) {
            if (ien_rows == 0) fail(SFEM_ERR_BOUNDS, "assemble_system: BoundsError at element 1");
            for (int a = 0; a < nen; ++a) {
                const int64_t v = ien[(int64_t)a * ien_rows];
                if (v < 1 || v > n_nodes) fail(SFEM_ERR_BOUNDS, "assemble_system: BoundsError at element 1");
            }
            if (sstep == 3 && first_gp_det(ndim, function_class, nen, node_list, ien, ien_rows) == 0.0)
                fail(SFEM_ERR_SINGULAR, "assemble_system: SingularException at element 1");
            fail(scode, scode == SFEM_ERR_DIMENSION ? "assemble_system: DimensionMismatch at element 1"
                                                    : "assemble_system: UndefVarError at element 1");
        }
        if (function_class == SFEM_Q2 && n_dof >= 2) {
            // COO sized for 2^ndim nodes (:143) but Q2 Ke has 9 nodes: the
            // scatter runs past the end at element e*.
            const int64_t len = n_elem * (int64_t)(n_dof << ndim) * (n_dof << ndim);
            const int64_t ke = (int64_t)nen * n_dof * nen * n_dof;
            const int64_t estar = len / ke;  // first element whose scatter passes len
            if (estar * 8 + 4 < m->pre_key) { m->pre_code = SFEM_ERR_BOUNDS; m->pre_key = estar * 8 + 4; }
        }
        m->coords.reset(ctx, (size_t)(ndim * n_nodes));
        h2d(ctx, m->coords.get(), node_list, sizeof(double) * ndim * n_nodes);
        {
            DevBuf<int64_t> raw(ctx, (size_t)(n_used * nen));
            if (n_used == ien_rows) {
                h2d(ctx, raw.get(), ien, sizeof(int64_t) * n_used * nen);
            } else {
                for (int a = 0; a < nen; ++a)
                    h2d(ctx, raw.get() + (int64_t)a * n_used, ien + (int64_t)a * ien_rows, sizeof(int64_t) * n_used);
            }
            m->ien.reset(ctx, (size_t)(n_used * nen));
            hipLaunchKernelGGL(convert_ien_kernel, dim3(grid_for(n_used * nen, 256)), dim3(256), 0, ctx->stream,
                               raw.get(), n_used, nen, n_used, n_nodes, m->ien.get());
            SFEM_LAUNCH_CHECK();
            sync(ctx);  // raw is released below; keep the copy complete
        }
        if (n_dof >= 2 && id_rows > 0) {
            DevBuf<int64_t> raw(ctx, (size_t)(id_rows * n_dof));
            h2d(ctx, raw.get(), id, sizeof(int64_t) * id_rows * n_dof);
            m->dof.reset(ctx, (size_t)(id_rows * n_dof));
            hipLaunchKernelGGL(convert_id_kernel, dim3(grid_for(id_rows * n_dof, 256)), dim3(256), 0, ctx->stream,
                               raw.get(), id_rows, n_dof, m->dof.get());
            SFEM_LAUNCH_CHECK();
            sync(ctx);
        }
        m->young.reset(ctx, (size_t)n_young);
        h2d(ctx, m->young.get(), young, sizeof(double) * n_young);
        sync(ctx);
    });
    if (st) { delete m; return st; }
    m->ctx = ctx;
    ctx_retain(ctx);
    *out = m;
    return SFEM_OK;
}

int sfem_mesh_set_grid(sfem_ctx* ctx, sfem_mesh* m, int ndim, const int64_t* ne) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!m) { ctx->last_error = "mesh_set_grid: NULL mesh"; return SFEM_ERR_ARGUMENT; }
    return guard(ctx, [&] { mesh_set_grid(ctx, m, ndim, ne); });
}

void sfem_mesh_destroy(sfem_mesh* m) {
    if (!m) return;
    sfem_ctx* ctx = m->ctx;
    delete m;
    ctx_release(ctx);
}

int sfem_assemble_mesh(sfem_ctx* ctx, const sfem_mesh* m, int flags, sfem_csr** out) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!m || !out) { ctx->last_error = "NULL mesh or out"; return SFEM_ERR_ARGUMENT; }
    if (m->ctx != ctx && m->n_elem) { ctx->last_error = "mesh belongs to another context"; return SFEM_ERR_STATE; }
    *out = nullptr;
    sfem_csr* K = new (std::nothrow) sfem_csr();
    if (!K) return SFEM_ERR_MEMORY;
    int st = guard(ctx, [&] {
        if (m->n_elem == 0) {
            if (m->pre_code) fail(m->pre_code, "assemble_system: BoundsError at element 1");
            K->ctx = ctx;
            K->bcolptr.reset(ctx, 1);
            memset_d(ctx, K->bcolptr.get(), 0, sizeof(int64_t));
            sync(ctx);
            return;
        }
        assemble(ctx, m, flags, K);
    });
    if (st) { delete K; return st; }
    K->ctx = ctx;
    ctx_retain(ctx);
    *out = K;
    return SFEM_OK;
}

int sfem_mesh_detect_grid(sfem_ctx* ctx, sfem_mesh* m, int* ndim, int64_t* ne) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!m) { ctx->last_error = "mesh_detect_grid: NULL mesh"; return SFEM_ERR_ARGUMENT; }
    return guard(ctx, [&] {
        int64_t d[3] = {0, 0, 0};
        const bool found = mesh_detect_grid(ctx, m, d);
        if (ndim) *ndim = found ? m->ndim : 0;
        if (ne)
            for (int k = 0; k < 3; ++k) ne[k] = found ? d[k] : 0;
    });
}

int sfem_assemble(sfem_ctx* ctx, int ndim, int function_class, int n_dof, const double* node_list,
                  int64_t n_nodes, const int64_t* ien, int64_t ien_rows, int nen, int64_t n_elem,
                  const int64_t* id, int64_t id_rows, const double* young, int64_t n_young, double nu,
                  sfem_csr** out) {
    sfem_mesh* m = nullptr;
    int st = sfem_mesh_create(ctx, ndim, function_class, n_dof, node_list, n_nodes, ien, ien_rows, nen, n_elem,
                              id, id_rows, young, n_young, nu, &m);
    if (st) return st;
    // a mesh in the meshgrid numbering is assembled from the grid (bitwise the
    // same K, no sort) and K keeps the element scales the multigrid reads
    st = sfem_mesh_detect_grid(ctx, m, nullptr, nullptr);
    if (!st) st = sfem_assemble_mesh(ctx, m, SFEM_ASM_DEFAULT, out);
    sfem_mesh_destroy(m);
    return st;
}

int sfem_csr_info(const sfem_csr* K, int64_t* m, int64_t* n, int64_t* nnz, int* block) {
    if (!K) return SFEM_ERR_ARGUMENT;
    if (m) *m = K->m;
    if (n) *n = K->n;
    if (nnz) *nnz = K->nnz;
    if (block) *block = K->bs;
    return SFEM_OK;
}

int sfem_csr_grid(const sfem_csr* K, int* ndim, int64_t* ne) {
    if (!K) return SFEM_ERR_ARGUMENT;
    if (ndim) *ndim = K->grid_nd;
    if (ne)
        for (int d = 0; d < 3; ++d) ne[d] = d < K->grid_nd ? K->grid_ne[d] : 0;
    return SFEM_OK;
}

int sfem_csr_grid_values(const sfem_csr* K, int* from_scales, double* ke_dev) {
    if (!K) return SFEM_ERR_ARGUMENT;
    if (from_scales) *from_scales = K->from_scales ? 1 : 0;
    if (ke_dev) *ke_dev = K->ke_dev;
    return SFEM_OK;
}

int sfem_csr_stage_ms(const sfem_csr* K, double* ms) {
    if (!K || !ms) return SFEM_ERR_ARGUMENT;
    ms[0] = K->t_validate; ms[1] = K->t_element; ms[2] = K->t_keys; ms[3] = K->t_sort;
    ms[4] = K->t_reduce; ms[5] = K->t_total;
    return SFEM_OK;
}

int sfem_export_csc1(sfem_ctx* ctx, const sfem_csr* K, int64_t* colptr, int64_t* rowval, double* nzval) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!K) { ctx->last_error = "NULL matrix"; return SFEM_ERR_ARGUMENT; }
    return guard(ctx, [&] {
        if (K->n == 0) {
            if (colptr) colptr[0] = 1;
            return;
        }
        export_csc1(ctx, K, colptr, rowval, nzval);
    });
}

void sfem_csr_destroy(sfem_csr* K) {
    if (!K) return;
    if (K->refs.fetch_sub(1) != 1) return;  // a system eliminated from K still reads it
    sfem_ctx* ctx = K->ctx;
    delete K;
    ctx_release(ctx);
}

int sfem_csr_from_csc1(sfem_ctx* ctx, int64_t m, int64_t n, const int64_t* colptr, const int64_t* rowval,
                       const double* nzval, sfem_csr** out) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!colptr || !out || m < 0 || n < 0) { ctx->last_error = "csr_from_csc1: bad arguments"; return SFEM_ERR_ARGUMENT; }
    *out = nullptr;
    sfem_csr* K = new (std::nothrow) sfem_csr();
    if (!K) return SFEM_ERR_MEMORY;
    int st = guard(ctx, [&] { upload_csc1(ctx, m, n, colptr, rowval, nzval, K); });
    if (st) { delete K; return st; }
    K->ctx = ctx;
    ctx_retain(ctx);
    *out = K;
    return SFEM_OK;
}

// ---------------------------------------------------------------- solve --
int sfem_bc_create(sfem_ctx* ctx, int64_t n, const double* f, const uint8_t* fixed, const double* q_d,
                   sfem_bc** out) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!out || n < 0) { ctx->last_error = "bc_create: bad arguments"; return SFEM_ERR_ARGUMENT; }
    *out = nullptr;
    sfem_bc* bc = new (std::nothrow) sfem_bc();
    if (!bc) return SFEM_ERR_MEMORY;
    int st = guard(ctx, [&] {
        bc->ctx = ctx;
        bc->n = n;
        bc->f.reset(ctx, (size_t)n);
        bc->fixed.reset(ctx, (size_t)n);
        bc->qd.reset(ctx, (size_t)n);
        std::vector<double> qd0;
        if (f) h2d(ctx, bc->f.get(), f, sizeof(double) * n);
        else memset_d(ctx, bc->f.get(), 0, sizeof(double) * n);
        if (fixed) h2d(ctx, bc->fixed.get(), fixed, (size_t)n);
        else memset_d(ctx, bc->fixed.get(), 0, (size_t)n);
        bool any = false;
        if (q_d && fixed) {  // prescribed values only on fixed dofs
            qd0.assign(q_d, q_d + n);
            for (int64_t i = 0; i < n; ++i) {
                if (!fixed[i]) qd0[i] = 0.0;
                any = any || qd0[i] != 0.0;
            }
            h2d(ctx, bc->qd.get(), qd0.data(), sizeof(double) * n);
        } else {
            memset_d(ctx, bc->qd.get(), 0, sizeof(double) * n);
        }
        bc->qd_zero = !any;
        sync(ctx);
    });
    if (st) { delete bc; return st; }
    ctx_retain(ctx);
    *out = bc;
    return SFEM_OK;
}

void sfem_bc_destroy(sfem_bc* bc) {
    if (!bc) return;
    sfem_ctx* ctx = bc->ctx;
    delete bc;
    ctx_release(ctx);
}

int sfem_eliminate(sfem_ctx* ctx, const sfem_csr* K, const sfem_bc* bc, sfem_system** out) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!K || !bc || !out) { ctx->last_error = "eliminate: NULL argument"; return SFEM_ERR_ARGUMENT; }
    *out = nullptr;
    sfem_system* S = new (std::nothrow) sfem_system();
    if (!S) return SFEM_ERR_MEMORY;
    int st = guard(ctx, [&] { eliminate(ctx, K, bc, S); });
    if (st) { delete S; return st; }
    S->ctx = ctx;
    ctx_retain(ctx);
    K->refs.fetch_add(1);  // dropped by sfem_system_destroy
    *out = S;
    return SFEM_OK;
}

int sfem_system_info(const sfem_system* S, int64_t* n_free, int64_t* nnz_ff, int* block) {
    if (!S) return SFEM_ERR_ARGUMENT;
    if (n_free) *n_free = S->nf;
    if (block) *block = S->has_block ? S->bs : 1;
    if (!nnz_ff) return SFEM_OK;
    // the row counts are formed on first request (the CSR path needs them too)
    return guard(S->ctx, [&] {
        ensure_rowptr(S->ctx, const_cast<sfem_system*>(S));
        *nnz_ff = S->nnz;
    });
}

int sfem_system_eliminate_ms(const sfem_system* S, double* ms) {
    if (!S || !ms) return SFEM_ERR_ARGUMENT;
    *ms = S->t_eliminate_ms;
    return SFEM_OK;
}

void sfem_system_destroy(sfem_system* S) {
    if (!S) return;
    sfem_ctx* ctx = S->ctx;
    sfem_csr* K = const_cast<sfem_csr*>(S->K);
    delete S;
    sfem_csr_destroy(K);  // the reference sfem_eliminate took
    ctx_release(ctx);
}

int sfem_system_solve_noalloc(sfem_ctx* ctx, sfem_system* S, const sfem_solve_opts* opts, sfem_solve_stats* stats) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!S || !opts || !stats) { ctx->last_error = "solve_noalloc: NULL argument"; return SFEM_ERR_ARGUMENT; }
    if (S->ctx != ctx) { ctx->last_error = "solve_noalloc: system belongs to another context"; return SFEM_ERR_STATE; }
    return guard(ctx, [&] { system_solve_noalloc(ctx, S, opts, stats); });
}

void sfem_solve_opts_default(sfem_solve_opts* o) {
    if (!o) return;
    o->precision = SFEM_PREC_FP64;
    o->format = SFEM_FMT_AUTO;
    o->rtol = 1e-10;
    o->max_iter = 200000;
    o->reliable_delta = 0.1;
    o->check_every = 64;
    o->time_kernels = 0;
    o->x0_zero = 1;
    o->preconditioner = SFEM_PC_JACOBI;
}

int sfem_system_solve(sfem_ctx* ctx, sfem_system* S, const sfem_solve_opts* opts, sfem_solve_stats* stats) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!S) { ctx->last_error = "solve: NULL system"; return SFEM_ERR_ARGUMENT; }
    sfem_solve_opts o;
    sfem_solve_opts_default(&o);
    if (opts) o = *opts;
    sfem_solve_stats local{};
    return guard(ctx, [&] { system_solve(ctx, S, &o, stats ? stats : &local); });
}

void sfem_mg_opts_default(sfem_mg_opts* o) {
    if (o) mg_opts_default(o);
}

int sfem_system_set_grid(sfem_ctx* ctx, sfem_system* S, int ndim, const int64_t* dims, const sfem_mg_opts* opts) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!S) { ctx->last_error = "set_grid: NULL system"; return SFEM_ERR_ARGUMENT; }
    return guard(ctx, [&] { system_set_grid(ctx, S, ndim, dims, opts); });
}

int sfem_system_mg_info(const sfem_system* S, sfem_mg_info* info) {
    if (!S || !info) return SFEM_ERR_ARGUMENT;
    return guard(S->ctx, [&] { mg_info(S, info); });
}

int sfem_system_mg_stencil(sfem_ctx* ctx, const sfem_system* S, int level, double* stencil, uint8_t* fixed) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!S) { ctx->last_error = "mg_stencil: NULL system"; return SFEM_ERR_ARGUMENT; }
    return guard(ctx, [&] { mg_stencil(ctx, S, level, stencil, fixed); });
}

int sfem_system_solution(sfem_ctx* ctx, const sfem_system* S, double* q, double* reactions) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!S) { ctx->last_error = "solution: NULL system"; return SFEM_ERR_ARGUMENT; }
    return guard(ctx, [&] { system_solution(ctx, S, q, reactions); });
}

int sfem_solve(sfem_ctx* ctx, const sfem_csr* K, const double* f, const uint8_t* fixed, const double* q_d,
               const sfem_solve_opts* opts, double* q, double* reactions, sfem_solve_stats* stats) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!K) { ctx->last_error = "solve: NULL matrix"; return SFEM_ERR_ARGUMENT; }
    sfem_bc* bc = nullptr;
    int st = sfem_bc_create(ctx, K->n, f, fixed, q_d, &bc);
    if (st) return st;
    sfem_system* S = nullptr;
    st = sfem_eliminate(ctx, K, bc, &S);
    sfem_bc_destroy(bc);
    if (st) return st;
    st = sfem_system_solve(ctx, S, opts, stats);
    if (st == SFEM_OK || st == SFEM_ERR_NOT_CONVERGED) {
        const int st2 = sfem_system_solution(ctx, S, q, reactions);
        if (st2) st = st2;
    }
    sfem_system_destroy(S);
    return st;
}

int sfem_system_spmv_bench(sfem_ctx* ctx, sfem_system* S, int precision, int format, int reps, double* avg_ms,
                           double* bytes) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!S || !avg_ms || !bytes) { ctx->last_error = "spmv_bench: NULL argument"; return SFEM_ERR_ARGUMENT; }
    return guard(ctx, [&] { system_spmv_bench(ctx, S, precision, format, reps, avg_ms, bytes); });
}

int sfem_stream_copy_bench(sfem_ctx* ctx, int64_t bytes, int reps, double* avg_ms) {
    if (!ctx) return SFEM_ERR_ARGUMENT;
    if (!avg_ms) return SFEM_ERR_ARGUMENT;
    return guard(ctx, [&] { stream_copy_bench(ctx, bytes, reps, avg_ms); });
}

}  // extern "C"
