/*
 * sfem_oracle.c -- CPU restatement of smearFEM's assemble-and-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU library is
 * compared against (and the CPU baseline bench.py times).  Nothing in the
 * product library links, loads or calls it.
 *
 * It follows the reference's algorithm step by step, in the reference's
 * order, citing soshalaw/smearFEM.jl file:line:
 *   gaussian_quadrature      src/fem.jl:21-31
 *   basis_function           src/fem.jl:48-114
 *   assemble_system          src/fem.jl:135-256 (element loop :179, gp loop :183,
 *                            Jacobian :192-196, Laplace :198-208, B/D/Ke :209-233,
 *                            COO scatter :236-249)
 *   sparse(E,J,V)            src/fem.jl:253 -> Julia stdlib SparseArrays.sparse!
 *                            (version unpinned: no Manifest in the reference):
 *                            stable counting sort by row, duplicates combined
 *                            with + in input order, explicit zeros kept,
 *                            transpose to CSC; dims = max index.
 *   apply_boundary_conditions examples/vector3D.jl:175-264 (surface penalty b)
 *   Dirichlet elimination    examples/vector3D.jl:151,170,315-316,320-322
 *   dense solve              examples/vector3D.jl:318 (inv(Matrix(K_free)))
 * plus FP64 Jacobi-PCG (no reference counterpart; CPU baseline for sizes a
 * dense or banded direct solve cannot reach).
 *
 * Status codes are those of include/smearfem.h (SFEM_ERR_*), reproduced here
 * as plain integers so the oracle builds with gcc alone.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_OK 0
#define OR_ERR_ARGUMENT -1
#define OR_ERR_BOUNDS -2
#define OR_ERR_DIMENSION -3
#define OR_ERR_SINGULAR -4
#define OR_ERR_UNDEFVAR -5
#define OR_ERR_NOT_CONVERGED -6
#define OR_ERR_MEMORY -7

/* ---- src/fem.jl:21-31 -------------------------------------------------- */
int or_gaussian_quadrature(double a, double b, int n, double* xi, double* w) {
    if (n == 2) {
        xi[0] = -(b - a) / (2 * sqrt(3.0)) + (b + a) / 2;
        xi[1] = (b - a) / (2 * sqrt(3.0)) + (b + a) / 2;
        w[0] = (b - a) / 2;
        w[1] = (b - a) / 2;
        return OR_OK;
    }
    if (n == 3) {
        xi[0] = -(b - a) / (2 * sqrt(5.0 / 3.0)) + (b + a) / 2;
        xi[1] = 0;
        xi[2] = (b - a) / (2 * sqrt(5.0 / 3.0)) + (b + a) / 2;
        w[0] = (b - a) / 2 * 5 / 9;
        w[1] = (b - a) / 2 * 8 / 9;
        w[2] = (b - a) / 2 * 5 / 9;
        return OR_OK;
    }
    return OR_ERR_UNDEFVAR; /* no else branch: ξ undefined at :30 */
}

/* ---- src/fem.jl:48-114 ------------------------------------------------- *
 * dN is column-major nfun x ncol (Julia layout).  The 1D ΔN is the 1x2 row
 * [-0.5 0.5] (:75), reported as nrow=1, ncol=2. */
int or_basis_function(int nargs, double x, double y, double z, int cls, double* N, double* dN,
                      int* nfun, int* nrow, int* ncol) {
    if (cls == 1) {
        if (nargs >= 3) {
            const double s[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                    {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
            for (int a = 0; a < 8; ++a) {
                double fx = 1 + s[a][0] * x, fy = 1 + s[a][1] * y, fz = 1 + s[a][2] * z;
                N[a] = fx * fy * fz / 8;
                dN[a + 0 * 8] = s[a][0] * fy * fz / 8;
                dN[a + 1 * 8] = s[a][1] * fx * fz / 8;
                dN[a + 2 * 8] = s[a][2] * fx * fy / 8;
            }
            *nfun = 8; *nrow = 8; *ncol = 3;
            return OR_OK;
        }
        if (nargs == 2) {
            N[0] = (1 - x) * (1 - y) / 4;
            N[1] = (x + 1) * (1 - y) / 4;
            N[2] = (1 + x) * (y + 1) / 4;
            N[3] = (1 - x) * (1 + y) / 4;
            dN[0] = -(1 - y) / 4; dN[1] = (1 - y) / 4; dN[2] = (y + 1) / 4; dN[3] = -(1 + y) / 4;
            dN[4] = -(1 - x) / 4; dN[5] = -(x + 1) / 4; dN[6] = (1 + x) / 4; dN[7] = (1 - x) / 4;
            *nfun = 4; *nrow = 4; *ncol = 2;
            return OR_OK;
        }
        N[0] = 0.5 - 0.5 * x;
        N[1] = 0.5 + 0.5 * x;
        dN[0] = -0.5; dN[1] = 0.5;
        *nfun = 2; *nrow = 1; *ncol = 2;
        return OR_OK;
    }
    if (cls == 2 && nargs >= 2) {
        N[0] = (1 - x) * x * (1 - y) * y / 4;
        N[1] = -x * (1 + x) * (1 - y) * y / 4;
        N[2] = x * (1 + x) * y * (1 + y) / 4;
        N[3] = -(1 - x) * x * y * (1 + y) / 4;
        N[4] = -(1 - x) * (1 + x) * (1 - y) * y / 2;
        N[5] = x * (1 + x) * (1 - y) * (1 + y) / 2;
        N[6] = (1 - x) * (1 + x) * y * (1 + y) / 2;
        N[7] = -(1 - x) * x * (1 - y) * (1 + y) / 2;
        N[8] = (1 - x) * (1 + x) * (1 - y) * (1 + y);
        double* Nx = dN;
        double* Ny = dN + 9;
        Nx[0] = (1 - 2 * x) * (1 - y) * y / 4;
        Nx[1] = -(1 + 2 * x) * (1 - y) * y / 4;
        Nx[2] = (1 + 2 * x) * y * (1 + y) / 4;
        Nx[3] = -(1 - 2 * x) * y * (1 + y) / 4;
        Nx[4] = x * (1 - y) * y;
        Nx[5] = (1 + 2 * x) * (1 - y) * (1 + y) / 2;
        Nx[6] = -x * y * (1 + y);
        Nx[7] = -(1 - 2 * x) * (1 - y) * (1 + y) / 2;
        Nx[8] = -2 * x * (1 - y) * (1 + y);
        Ny[0] = (1 - x) * x * (1 - 2 * y) / 4;
        Ny[1] = -x * (1 + x) * (1 - 2 * y) / 4;
        Ny[2] = x * (1 + x) * (1 + 2 * y) / 4;
        Ny[3] = -(1 - x) * x * (1 + 2 * y) / 4;
        Ny[4] = -(1 - x) * (1 + x) * (1 - 2 * y) / 2;
        Ny[5] = -x * (1 + x) * y;
        Ny[6] = (1 - x) * (1 + x) * (1 + 2 * y) / 2;
        Ny[7] = (1 - x) * x * y;
        Ny[8] = -(1 - x) * (1 + x) * 2 * y;
        *nfun = 9; *nrow = 9; *ncol = 2;
        return OR_OK;
    }
    return OR_ERR_UNDEFVAR; /* N never assigned: UndefVarError at :113 */
}

/* LU with partial pivoting of a small dense matrix (n<=3), as LAPACK getrf
 * does for det (src/fem.jl:194) and inv (src/fem.jl:195).  Returns the index
 * (1-based) of the first exactly-zero pivot, 0 if none. */
static int lu_small(int n, double* A /* row-major n x n */, int* piv, double* detv) {
    int info = 0;
    double sgn = 1.0;
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = fabs(A[k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (fabs(A[i * n + k]) > best) { best = fabs(A[i * n + k]); p = i; }
        piv[k] = p;
        if (p != k) {
            sgn = -sgn;
            for (int j = 0; j < n; ++j) { double t = A[k * n + j]; A[k * n + j] = A[p * n + j]; A[p * n + j] = t; }
        }
        if (A[k * n + k] == 0.0) { if (!info) info = k + 1; continue; }
        for (int i = k + 1; i < n; ++i) {
            A[i * n + k] /= A[k * n + k];
            for (int j = k + 1; j < n; ++j) A[i * n + j] -= A[i * n + k] * A[k * n + j];
        }
    }
    double d = sgn;
    for (int k = 0; k < n; ++k) d *= A[k * n + k];
    *detv = d;
    return info;
}

/* inverse from the LU factors (columns of inv solve LU x = P e_j) */
static void lu_inverse(int n, const double* LU, const int* piv, double* inv /* row-major */) {
    for (int j = 0; j < n; ++j) {
        double x[3] = {0, 0, 0};
        x[j] = 1.0;
        for (int k = 0; k < n; ++k) { double t = x[k]; x[k] = x[piv[k]]; x[piv[k]] = t; }
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < i; ++k) x[i] -= LU[i * n + k] * x[k];
        for (int i = n - 1; i >= 0; --i) {
            for (int k = i + 1; k < n; ++k) x[i] -= LU[i * n + k] * x[k];
            x[i] /= LU[i * n + i];
        }
        for (int i = 0; i < n; ++i) inv[i * n + j] = x[i];
    }
}

/* quadrature tables, src/fem.jl:149-177 (2-point rule per direction, CCW
 * then bottom/top order) */
static int gauss_table(int ndim, double gx[8], double gy[8], double gz[8], double gw[8]) {
    double xi[2], w[2];
    or_gaussian_quadrature(-1, 1, 2, xi, w);
    if (ndim == 2) {
        const int ix[4] = {0, 1, 1, 0}, iy[4] = {0, 0, 1, 1};
        for (int g = 0; g < 4; ++g) { gx[g] = xi[ix[g]]; gy[g] = xi[iy[g]]; gz[g] = 0; gw[g] = w[ix[g]] * w[iy[g]]; }
        return 4;
    }
    const int ix[8] = {0, 1, 1, 0, 0, 1, 1, 0}, iy[8] = {0, 0, 1, 1, 0, 0, 1, 1}, iz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    for (int g = 0; g < 8; ++g) {
        gx[g] = xi[ix[g]]; gy[g] = xi[iy[g]]; gz[g] = xi[iz[g]];
        gw[g] = w[ix[g]] * w[iy[g]] * w[iz[g]];
    }
    return 8;
}

/* Per-gp element kernel, src/fem.jl:184-233.  coords: ndim x nen column-major
 * (x of node a at coords[d + ndim*a]).  Writes Ke_gp (L x L, row-major, L =
 * nen*ndof; already multiplied by w) for gp g.  Returns OR_OK/OR_ERR_*. */
static int element_gp(int ndim, int cls, int ndof, int nen, const double* coords, double E, double nu,
                      double xg, double yg, double zg, double wg, double* Ke) {
    double N[9], dN[27];
    int nfun, nrow, ncol;
    int st = or_basis_function(ndim, xg, yg, zg, cls, N, dN, &nfun, &nrow, &ncol);
    if (st) return st;
    if (ndim == 1 || nrow != nen || ncol != ndim) return OR_ERR_DIMENSION; /* coords*ΔN at :192 */
    double J[9], LU[9], invJ[9];
    int piv[3];
    for (int i = 0; i < ndim; ++i)
        for (int j = 0; j < ndim; ++j) {
            double s = 0;
            for (int a = 0; a < nen; ++a) s += coords[i + ndim * a] * dN[a + nrow * j];
            J[i * ndim + j] = s;
        }
    memcpy(LU, J, sizeof(double) * ndim * ndim);
    double detJ;
    int info = lu_small(ndim, LU, piv, &detJ);
    double w = wg * fabs(detJ);                  /* :194 */
    if (info) return OR_ERR_SINGULAR;            /* inv(Jac), :195 */
    lu_inverse(ndim, LU, piv, invJ);
    double dNdX[27];                             /* nen x ndim, :196 */
    for (int a = 0; a < nen; ++a)
        for (int k = 0; k < ndim; ++k) {
            double s = 0;
            for (int j = 0; j < ndim; ++j) s += dN[a + nrow * j] * invJ[j * ndim + k];
            dNdX[a * ndim + k] = s;
        }
    if (ndof == 1) { /* :198-208 */
        for (int i = 0; i < nen; ++i)
            for (int j = 0; j < nen; ++j) {
                double s = 0;
                for (int k = 0; k < ndim; ++k) s += dNdX[i * ndim + k] * dNdX[j * ndim + k];
                Ke[i * nen + j] = w * s;
            }
        return OR_OK;
    }
    int L = ndim * nen, nv;
    double B[6 * 27], D[36];
    memset(B, 0, sizeof(B));
    memset(D, 0, sizeof(D));
    if (ndof == 2) { /* :210-217, plane stress */
        if (ndim != 2) return OR_ERR_DIMENSION;
        nv = 3;
        for (int a = 0; a < nen; ++a) {
            B[0 * L + 2 * a] = dNdX[a * 2 + 0];
            B[1 * L + 2 * a + 1] = dNdX[a * 2 + 1];
            B[2 * L + 2 * a] = dNdX[a * 2 + 1];
            B[2 * L + 2 * a + 1] = dNdX[a * 2 + 0];
        }
        double c = E / (1 - nu * nu);
        D[0 * 6 + 0] = c; D[0 * 6 + 1] = nu * c;
        D[1 * 6 + 0] = nu * c; D[1 * 6 + 1] = c;
        D[2 * 6 + 2] = E / (2 * (1 + nu));
    } else if (ndof == 3) { /* :218-230 */
        if (ndim != 3) return OR_ERR_DIMENSION;
        nv = 6;
        for (int a = 0; a < nen; ++a) {
            double gx = dNdX[a * 3], gy = dNdX[a * 3 + 1], gz = dNdX[a * 3 + 2];
            B[0 * L + 3 * a] = gx;
            B[1 * L + 3 * a + 1] = gy;
            B[2 * L + 3 * a + 2] = gz;
            B[3 * L + 3 * a + 1] = gz; B[3 * L + 3 * a + 2] = gy;
            B[4 * L + 3 * a] = gz;     B[4 * L + 3 * a + 2] = gx;
            B[5 * L + 3 * a] = gy;     B[5 * L + 3 * a + 1] = gx;
        }
        double c = E / ((1 + nu) * (1 - 2 * nu));
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) D[i * 6 + j] = (i == j ? 1 - nu : nu) * c;
        for (int i = 3; i < 6; ++i) D[i * 6 + i] = (1 - 2 * nu) / 2 * c;
    } else {
        return OR_ERR_UNDEFVAR; /* B never assigned for nDof >= 4 */
    }
    /* Ke = ((B' * D) * B) * w, :233 */
    double BtD[27 * 6];
    for (int i = 0; i < L; ++i)
        for (int j = 0; j < nv; ++j) {
            double s = 0;
            for (int k = 0; k < nv; ++k) s += B[k * L + i] * D[k * 6 + j];
            BtD[i * 6 + j] = s;
        }
    for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j) {
            double s = 0;
            for (int k = 0; k < nv; ++k) s += BtD[i * 6 + k] * B[k * L + j];
            Ke[i * L + j] = s * w;
        }
    return OR_OK;
}

/* Element stiffness summed over the gp loop in gp order (V[inz] += Ke_gp),
 * for fixture tests.  Ke out: L x L row-major. */
int or_element_stiffness(int ndim, int cls, int ndof, int nen, const double* coords, double E, double nu,
                         double* Ke) {
    double gx[8], gy[8], gz[8], gw[8];
    if (ndim < 2 || ndim > 3) return OR_ERR_DIMENSION;
    int ngp = gauss_table(ndim, gx, gy, gz, gw);
    int L = nen * ndof;
    double Kg[27 * 27];
    memset(Ke, 0, sizeof(double) * L * L);
    for (int g = 0; g < ngp; ++g) {
        int st = element_gp(ndim, cls, ndof, nen, coords, E, nu, gx[g], gy[g], gz[g], gw[g], Kg);
        if (st) return st;
        for (int i = 0; i < L * L; ++i) Ke[i] += Kg[i];
    }
    return OR_OK;
}

/* COO length the reference allocates, src/fem.jl:138-146 */
int64_t or_coo_length(int ndim, int ndof, int nen, int64_t n_elem) {
    if (ndof == 1) return n_elem * (int64_t)nen * nen;
    int64_t m = (int64_t)ndof << ndim; /* size(ID,2)*2^ndim */
    return n_elem * m * m;
}

/* assemble_system up to (but excluding) sparse(): fills E, J, V (1-based
 * indices, zero-initialised by the caller like zeros(...) at :139-145) in the
 * reference's element / gp / scatter order.  On error returns the status and
 * the 0-based element index where the reference would throw. */
int or_assemble_coo(int ndim, int cls, int ndof, const double* nodes, int64_t n_nodes,
                    const int64_t* ien, int64_t ien_rows, int nen, int64_t n_elem,
                    const int64_t* id, int64_t id_rows, const double* young, int64_t n_young,
                    double nu, int64_t* E, int64_t* J, double* V, int64_t* err_elem) {
    double gx[8], gy[8], gz[8], gw[8];
    *err_elem = -1;
    if (ndof >= 2 && !id) return OR_ERR_ARGUMENT;
    int ngp = (ndim == 2 || ndim == 3) ? gauss_table(ndim, gx, gy, gz, gw) : 2;
    int64_t ncoo = or_coo_length(ndim, ndof, nen, n_elem);
    double coords[27];
    double Kg[27 * 27];
    for (int64_t e = 0; e < n_elem; ++e) {
        *err_elem = e;
        if (e >= ien_rows) return OR_ERR_BOUNDS;               /* IEN[e,:] */
        for (int a = 0; a < nen; ++a) {                        /* NodeList[:,IEN[e,:]], :180 */
            int64_t nd = ien[a * ien_rows + e];
            if (nd < 1 || nd > n_nodes) return OR_ERR_BOUNDS;
            for (int d = 0; d < ndim; ++d) coords[d + ndim * a] = nodes[d + ndim * (nd - 1)];
        }
        double Ee = young[n_young == 1 ? 0 : e];
        for (int g = 0; g < ngp; ++g) {
            if (ndim != 2 && ndim != 3) return OR_ERR_DIMENSION;
            int st = element_gp(ndim, cls, ndof, nen, coords, Ee, nu, gx[g], gy[g], gz[g], gw[g], Kg);
            if (st) return st;
            if (ndof == 1) { /* :201-207 */
                for (int i = 0; i < nen; ++i)
                    for (int j = 0; j < nen; ++j) {
                        int64_t inz = (int64_t)nen * nen * e + (int64_t)nen * i + j;
                        E[inz] = ien[i * ien_rows + e];
                        J[inz] = ien[j * ien_rows + e];
                        V[inz] += Kg[i * nen + j];
                    }
            } else { /* :236-249 */
                int L = nen * ndof;
                for (int iN = 0; iN < nen; ++iN)
                    for (int jN = 0; jN < nen; ++jN)
                        for (int iD = 0; iD < ndof; ++iD)
                            for (int jD = 0; jD < ndof; ++jD) {
                                int i = iN * ndof + iD, j = jN * ndof + jD;
                                int64_t inz = (int64_t)L * L * e + (int64_t)iN * ndof * L + (int64_t)jN * ndof * ndof + iD * ndof + jD;
                                int64_t ni = ien[iN * ien_rows + e], nj = ien[jN * ien_rows + e];
                                if (ni > id_rows || nj > id_rows) return OR_ERR_BOUNDS;
                                if (inz >= ncoo) return OR_ERR_BOUNDS; /* Q2 with nDof >= 2 */
                                E[inz] = id[iD * id_rows + (ni - 1)];
                                J[inz] = id[jD * id_rows + (nj - 1)];
                                V[inz] += Kg[i * L + j];
                            }
            }
        }
    }
    *err_elem = -1;
    return OR_OK;
}

/* ---- examples/vector3D.jl:175-264 (apply_boundary_conditions) ---------- *
 * COO of the surface penalty b on the bottom and top faces, before sparse():
 * length 2 * n_faces * (ndof * nfn)^2 (:182-184), bottom faces first (inz_btm,
 * :235), then top faces (inz_top, :236).  Per face and Gauss point (x, y in
 * the order [ξ1 ξ2 ξ2 ξ1], [η1 η1 η2 η2], weights wξ*wη, :203-208):
 * N, ΔN = basis_function(x, y) (2D Q1), dxdξ = coords * ΔN (3x2, :219-220),
 * w = wgp * |dxdξ[:,1] x dxdξ[:,2]| (:222-223), be = M'M with
 * M[d, (a-1)*nDof + d] = N_a (:225-230), so be[(a,i),(b,j)] = N_a N_b if i == j
 * else 0; V[inz] += w * be[i,j] in Gauss-point order (:241-249).
 * Only ndim = 3, nfn = 4, nDof = 3 runs in the reference: ndim = 2 makes
 * coords*ΔN a DimensionMismatch (1x2 ΔN, :213-220), and nDof != 3 a
 * DimensionMismatch in M[1,1:nDof:end] = N (:227).  The norm is the 2-norm
 * sqrt(c1^2 + c2^2 + c3^2) (Julia's norm calls BLAS nrm2, which may scale;
 * the two agree to rounding). */
int or_surface_coo(const double* nodes, int64_t n_nodes, const int64_t* ien_btm, const int64_t* ien_top,
                   int64_t n_faces, int nfn, int ndim, int ndof, const int64_t* id, int64_t id_rows,
                   int64_t* E, int64_t* J, double* V, int64_t* err_face) {
    *err_face = -1;
    if (ndim != 3 || nfn != 4 || ndof != 3) return OR_ERR_DIMENSION;
    double xi[2], wx[2];
    or_gaussian_quadrature(-1, 1, 2, xi, wx);
    const double gxs[4] = {xi[0], xi[1], xi[1], xi[0]};
    const double gys[4] = {xi[0], xi[0], xi[1], xi[1]};
    const double gws[4] = {wx[0] * wx[0], wx[1] * wx[0], wx[1] * wx[1], wx[0] * wx[1]};
    const int L = nfn * ndof;          /* 12 */
    const int64_t LL = (int64_t)L * L; /* length(be) = 144 */
    for (int64_t e = 0; e < n_faces; ++e) {
        *err_face = e;
        double ct[12], cb[12]; /* coords_top / coords_btm, 3 x 4 column-major */
        for (int k = 0; k < nfn; ++k) {
            const int64_t nt = ien_top[k * n_faces + e], nb = ien_btm[k * n_faces + e];
            if (nt < 1 || nt > n_nodes || nb < 1 || nb > n_nodes) return OR_ERR_BOUNDS;
            for (int d = 0; d < 3; ++d) {
                ct[d + 3 * k] = nodes[d + 3 * (nt - 1)];
                cb[d + 3 * k] = nodes[d + 3 * (nb - 1)];
            }
        }
        for (int g = 0; g < 4; ++g) {
            double N[4], dN[8];
            int nf, nr, nc;
            or_basis_function(2, gxs[g], gys[g], 0, 1, N, dN, &nf, &nr, &nc);
            double tt[6], tb[6]; /* dxdξ, 3 x 2 column-major */
            for (int k = 0; k < 2; ++k)
                for (int d = 0; d < 3; ++d) {
                    double st = 0, sb = 0;
                    for (int q = 0; q < 4; ++q) {
                        st += ct[d + 3 * q] * dN[q + 4 * k];
                        sb += cb[d + 3 * q] * dN[q + 4 * k];
                    }
                    tt[d + 3 * k] = st;
                    tb[d + 3 * k] = sb;
                }
            double xt[3] = {tt[1] * tt[5] - tt[2] * tt[4], tt[2] * tt[3] - tt[0] * tt[5], tt[0] * tt[4] - tt[1] * tt[3]};
            double xb[3] = {tb[1] * tb[5] - tb[2] * tb[4], tb[2] * tb[3] - tb[0] * tb[5], tb[0] * tb[4] - tb[1] * tb[3]};
            const double w_top = gws[g] * sqrt(xt[0] * xt[0] + xt[1] * xt[1] + xt[2] * xt[2]);
            const double w_btm = gws[g] * sqrt(xb[0] * xb[0] + xb[1] * xb[1] + xb[2] * xb[2]);
            for (int iN = 0; iN < nfn; ++iN)
                for (int jN = 0; jN < nfn; ++jN)
                    for (int iD = 0; iD < ndof; ++iD)
                        for (int jD = 0; jD < ndof; ++jD) {
                            const double be = iD == jD ? N[iN] * N[jN] : 0.0;
                            const int64_t off = (int64_t)iN * ndof * L + (int64_t)jN * ndof * ndof + iD * ndof + jD;
                            const int64_t inz_btm = LL * e + off;
                            const int64_t inz_top = LL * n_faces + LL * e + off;
                            const int64_t ti = ien_top[iN * n_faces + e], tj = ien_top[jN * n_faces + e];
                            const int64_t bi = ien_btm[iN * n_faces + e], bj = ien_btm[jN * n_faces + e];
                            if (ti > id_rows || tj > id_rows || bi > id_rows || bj > id_rows) return OR_ERR_BOUNDS;
                            E[inz_top] = id[iD * id_rows + (ti - 1)];
                            J[inz_top] = id[jD * id_rows + (tj - 1)];
                            V[inz_top] += w_top * be;
                            E[inz_btm] = id[iD * id_rows + (bi - 1)];
                            J[inz_btm] = id[jD * id_rows + (bj - 1)];
                            V[inz_btm] += w_btm * be;
                        }
        }
    }
    *err_face = -1;
    return OR_OK;
}

/* dims of sparse(I,J,V): (max I, max J); ArgumentError for any index < 1 */
int or_sparse_dims(int64_t ncoo, const int64_t* I, const int64_t* J, int64_t* m, int64_t* n) {
    int64_t mi = 0, mj = 0;
    for (int64_t k = 0; k < ncoo; ++k) {
        if (I[k] < 1 || J[k] < 1) return OR_ERR_ARGUMENT;
        if (I[k] > mi) mi = I[k];
        if (J[k] > mj) mj = J[k];
    }
    *m = mi;
    *n = mj;
    return OR_OK;
}

/* sparse(I,J,V,m,n) with combine = + (SparseArrays.sparse!): counting sort
 * into rows (stable), combine repeats within a row in input order, then
 * counting sort by column into CSC.  colptr (n+1), rowval/nzval (>= nnz) are
 * 1-based on output.  Returns nnz via *nnz. */
int or_sparse_csc(int64_t ncoo, const int64_t* I, const int64_t* J, const double* V, int64_t m, int64_t n,
                  int64_t* colptr, int64_t* rowval, double* nzval, int64_t* nnz) {
    int64_t* rowptr = (int64_t*)calloc((size_t)m + 2, sizeof(int64_t));
    int64_t* ccol = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ncoo ? ncoo : 1));
    double* cval = (double*)malloc(sizeof(double) * (size_t)(ncoo ? ncoo : 1));
    int64_t* last = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    if (!rowptr || !ccol || !cval || !last) { free(rowptr); free(ccol); free(cval); free(last); return OR_ERR_MEMORY; }
    for (int64_t k = 0; k < ncoo; ++k) {
        if (I[k] < 1 || I[k] > m || J[k] < 1 || J[k] > n) { free(rowptr); free(ccol); free(cval); free(last); return OR_ERR_ARGUMENT; }
        rowptr[I[k] + 1]++;
    }
    for (int64_t i = 1; i <= m + 1; ++i) rowptr[i] += rowptr[i - 1];          /* rowptr[i] = start of row i (1-based rows at index i) */
    int64_t* wpos = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m + 2));
    memcpy(wpos, rowptr, sizeof(int64_t) * (size_t)(m + 2));
    for (int64_t k = 0; k < ncoo; ++k) {                                       /* stable scatter in input order */
        int64_t p = wpos[I[k]]++;
        ccol[p] = J[k];
        cval[p] = V[k];
    }
    for (int64_t j = 0; j <= n; ++j) last[j] = -1;
    int64_t* colcnt = (int64_t*)calloc((size_t)n + 2, sizeof(int64_t));
    int64_t w = 0;
    int64_t* newrow = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m + 2));
    for (int64_t i = 1; i <= m; ++i) {                                         /* combine repeats in order */
        int64_t rowstart = w;
        newrow[i] = w;
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
            int64_t j = ccol[k];
            if (last[j] < rowstart) {
                last[j] = w;
                ccol[w] = j;
                cval[w] = cval[k];
                ++w;
                colcnt[j]++;
            } else {
                cval[last[j]] = cval[last[j]] + cval[k];
            }
        }
    }
    newrow[m + 1] = w;
    colptr[0] = 1;
    for (int64_t j = 1; j <= n; ++j) colptr[j] = colptr[j - 1] + colcnt[j];
    int64_t* cpos = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    for (int64_t j = 1; j <= n; ++j) cpos[j] = colptr[j - 1] - 1;
    for (int64_t i = 1; i <= m; ++i)                                           /* rows ascend within columns */
        for (int64_t k = newrow[i]; k < newrow[i + 1]; ++k) {
            int64_t p = cpos[ccol[k]]++;
            rowval[p] = i;
            nzval[p] = cval[k];
        }
    *nnz = w;
    free(rowptr); free(ccol); free(cval); free(last); free(wpos); free(colcnt); free(newrow); free(cpos);
    return OR_OK;
}

/* ---- Dirichlet elimination, examples/vector3D.jl:151,170,315-316,320 ----
 * K given as 1-based CSC (n x n).  free_map[i] = new 0-based index of dof i
 * if free else -1.  Outputs K_ff as 0-based CSR (rows ascending columns) and
 * rhs = C'(f - K q_d).  Because K_free = C' K C picks entries K[free,free]
 * with one product by 1.0 each, it equals the submatrix exactly. */
int64_t or_eliminate_nnz(int64_t n, const int64_t* colptr, const int64_t* rowval, const int64_t* free_map) {
    int64_t c = 0;
    for (int64_t j = 0; j < n; ++j) {
        if (free_map[j] < 0) continue;
        for (int64_t k = colptr[j] - 1; k < colptr[j + 1] - 1; ++k)
            if (free_map[rowval[k] - 1] >= 0) ++c;
    }
    return c;
}

/* K_ff returned as CSC of the free block (0-based): since it is a principal
 * submatrix, CSC(K_ff) = CSR(K_ff') -- callers treat it as CSC. */
int or_eliminate(int64_t n, const int64_t* colptr, const int64_t* rowval, const double* nzval,
                 const int64_t* free_map, const double* f, const double* q_d,
                 int64_t* ff_colptr, int32_t* ff_rowval, double* ff_val, double* rhs) {
    /* rhs = C'(f - K q_d): full product first, as the example does (:320) */
    double* Kq = (double*)calloc((size_t)n, sizeof(double));
    if (!Kq) return OR_ERR_MEMORY;
    for (int64_t j = 0; j < n; ++j) {
        double qj = q_d ? q_d[j] : 0.0;
        if (qj == 0.0) continue;
        for (int64_t k = colptr[j] - 1; k < colptr[j + 1] - 1; ++k) Kq[rowval[k] - 1] += nzval[k] * qj;
    }
    int64_t c = 0, nf = 0;
    for (int64_t j = 0; j < n; ++j) {
        if (free_map[j] < 0) continue;
        ff_colptr[nf] = c;
        for (int64_t k = colptr[j] - 1; k < colptr[j + 1] - 1; ++k) {
            int64_t r = free_map[rowval[k] - 1];
            if (r >= 0) { ff_rowval[c] = (int32_t)r; ff_val[c] = nzval[k]; ++c; }
        }
        rhs[nf] = (f ? f[j] : 0.0) - Kq[j];
        ++nf;
    }
    ff_colptr[nf] = c;
    free(Kq);
    return OR_OK;
}

/* y = A x for A given as CSC of a symmetric matrix used as CSR */
void or_spmv(int64_t n, const int64_t* ptr, const int32_t* idx, const double* val, const double* x, double* y) {
    for (int64_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k) s += val[k] * x[idx[k]];
        y[i] = s;
    }
}

/* Dense Cholesky solve of the symmetric K_ff (the reference inverts the
 * dense K_free, examples/vector3D.jl:318; for an SPD matrix the Cholesky
 * solve is the stable equivalent).  A: n x n row-major, overwritten. */
int or_dense_spd_solve(int64_t n, double* A, const double* b, double* x) {
    for (int64_t j = 0; j < n; ++j) {
        double d = A[j * n + j];
        for (int64_t k = 0; k < j; ++k) d -= A[j * n + k] * A[j * n + k];
        if (!(d > 0.0)) return OR_ERR_SINGULAR;
        d = sqrt(d);
        A[j * n + j] = d;
        for (int64_t i = j + 1; i < n; ++i) {
            double s = A[i * n + j];
            for (int64_t k = 0; k < j; ++k) s -= A[i * n + k] * A[j * n + k];
            A[i * n + j] = s / d;
        }
    }
    for (int64_t i = 0; i < n; ++i) {
        double s = b[i];
        for (int64_t k = 0; k < i; ++k) s -= A[i * n + k] * x[k];
        x[i] = s / A[i * n + i];
    }
    for (int64_t i = n - 1; i >= 0; --i) {
        double s = x[i];
        for (int64_t k = i + 1; k < n; ++k) s -= A[k * n + i] * x[k];
        x[i] = s / A[i * n + i];
    }
    return OR_OK;
}

/* FP64 Jacobi-PCG (textbook Hestenes-Stiefel with M = diag(A)), single
 * thread.  Stops when ||b - A x|| <= rtol ||b|| (recurrence residual).
 * Returns iterations via *iters.  x: initial guess in, solution out. */
int or_pcg(int64_t n, const int64_t* ptr, const int32_t* idx, const double* val, const double* b,
           double* x, double rtol, int64_t max_iter, int64_t* iters, double* relres) {
    double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n), *p = malloc(sizeof(double) * n),
           *q = malloc(sizeof(double) * n), *dinv = malloc(sizeof(double) * n);
    if (!r || !z || !p || !q || !dinv) { free(r); free(z); free(p); free(q); free(dinv); return OR_ERR_MEMORY; }
    for (int64_t i = 0; i < n; ++i) {
        double d = 0.0;
        for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k) if (idx[k] == i) d = val[k];
        dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
    }
    or_spmv(n, ptr, idx, val, x, q);
    double bb = 0.0, rz = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        r[i] = b[i] - q[i];
        z[i] = dinv[i] * r[i];
        p[i] = z[i];
        rz += r[i] * z[i];
        bb += b[i] * b[i];
    }
    double bn = sqrt(bb), rr = 0.0;
    for (int64_t i = 0; i < n; ++i) rr += r[i] * r[i];
    int64_t it = 0;
    int st = OR_ERR_NOT_CONVERGED;
    if (bn == 0.0 || sqrt(rr) <= rtol * bn) st = OR_OK;
    while (st != OR_OK && it < max_iter) {
        or_spmv(n, ptr, idx, val, p, q);
        double pq = 0.0;
        for (int64_t i = 0; i < n; ++i) pq += p[i] * q[i];
        double alpha = rz / pq;
        double rz_new = 0.0;
        rr = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
            z[i] = dinv[i] * r[i];
            rz_new += r[i] * z[i];
            rr += r[i] * r[i];
        }
        ++it;
        if (sqrt(rr) <= rtol * bn) { st = OR_OK; break; }
        double beta = rz_new / rz;
        rz = rz_new;
        for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    *iters = it;
    *relres = bn > 0 ? sqrt(rr) / bn : 0.0;
    free(r); free(z); free(p); free(q); free(dinv);
    return st;
}

/* or_pcg with its loops split over nthreads OpenMP threads: the same
 * algorithm, used only to time an all-core CPU baseline (the summation order,
 * and so the last bits, depend on the thread count). */
static void spmv_mt(int64_t n, const int64_t* ptr, const int32_t* idx, const double* val, const double* x,
                    double* y, int nthreads) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k) s += val[k] * x[idx[k]];
        y[i] = s;
    }
}

int or_pcg_mt(int64_t n, const int64_t* ptr, const int32_t* idx, const double* val, const double* b,
              double* x, double rtol, int64_t max_iter, int nthreads, int64_t* iters, double* relres) {
    double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n), *p = malloc(sizeof(double) * n),
           *q = malloc(sizeof(double) * n), *dinv = malloc(sizeof(double) * n);
    if (!r || !z || !p || !q || !dinv) { free(r); free(z); free(p); free(q); free(dinv); return OR_ERR_MEMORY; }
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double d = 0.0;
        for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k) if (idx[k] == i) d = val[k];
        dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
    }
    spmv_mt(n, ptr, idx, val, x, q, nthreads);
    double bb = 0.0, rz = 0.0;
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : bb, rz)
    for (int64_t i = 0; i < n; ++i) {
        r[i] = b[i] - q[i];
        z[i] = dinv[i] * r[i];
        p[i] = z[i];
        rz += r[i] * z[i];
        bb += b[i] * b[i];
    }
    double bn = sqrt(bb), rr = 0.0;
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : rr)
    for (int64_t i = 0; i < n; ++i) rr += r[i] * r[i];
    int64_t it = 0;
    int st = OR_ERR_NOT_CONVERGED;
    if (bn == 0.0 || sqrt(rr) <= rtol * bn) st = OR_OK;
    while (st != OR_OK && it < max_iter) {
        spmv_mt(n, ptr, idx, val, p, q, nthreads);
        double pq = 0.0;
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : pq)
        for (int64_t i = 0; i < n; ++i) pq += p[i] * q[i];
        double alpha = rz / pq;
        double rz_new = 0.0;
        rr = 0.0;
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : rz_new, rr)
        for (int64_t i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
            z[i] = dinv[i] * r[i];
            rz_new += r[i] * z[i];
            rr += r[i] * r[i];
        }
        ++it;
        if (sqrt(rr) <= rtol * bn) { st = OR_OK; break; }
        double beta = rz_new / rz;
        rz = rz_new;
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    *iters = it;
    *relres = bn > 0 ? sqrt(rr) / bn : 0.0;
    free(r); free(z); free(p); free(q); free(dinv);
    return st;
}
