/*
 * Oracle weight computation (test infrastructure only).
 *
 * Restates init_rbffd_HH_rot_Ms and get_rbffd_DMs with their scalar kernels
 * (Advection/C_MPI_OMP_ACC/code/rbffd_DMs.c:55-328) and the builder's config-mode
 * method (PHS + tangent-plane polynomials on radius-scaled stencils, PHS Delta^k
 * system, 1D RBF-FD vertical stencils).  The dense solver is a textbook
 * right-looking LU with partial pivoting (first maximal |a_ik| as LAPACK idamax)
 * in place of MKL LAPACKE_dgesv (rbffd_DMs.c:240,243).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* rbffd_DMs.c:55-96.  v = (x, y, 1); if x^2+y^2 == 0 then beta = 0 (H = I, the
 * reference's south-pole case) else v2 = z>0 ? -(x^2+y^2)/(z+1) : z-1,
 * beta = 2 v2^2 / (x^2+y^2+v2^2), v = v/v2; H = I - beta v v^T. */
void orc_householder(double x, double y, double z, int fix_pole, double H[9])
{
    double dot_lon = x * x + y * y;
    double v[3] = {x, y, 1.0};
    double beta;
    if (dot_lon == 0.0) {
        beta = 0.0;
    } else {
        v[2] = z > 0.0 ? -dot_lon / (z + 1.0) : z - 1.0;
        beta = 2.0 * (v[2] * v[2]) / (dot_lon + (v[2] * v[2]));
        v[0] = v[0] / v[2];
        v[1] = v[1] / v[2];
        v[2] = v[2] / v[2];
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            H[3 * i + j] = (i == j ? 1.0 : 0.0) + (-beta * v[i]) * v[j];
    if (fix_pole && dot_lon == 0.0 && z < 0.0) {
        memset(H, 0, 9 * sizeof(double));
        H[0] = 1.0;
        H[4] = -1.0;
        H[8] = -1.0;
    }
}

/* rbffd_DMs.c:290-328 */
static double phs(double r, int k)
{
    if (k % 2 == 0)
        return r == 0.0 ? r : log(r) * pow(r, k);
    return pow(r, k);
}
static double d1_phs(double r, double ri, int k)
{
    if (k % 2 == 0)
        return r == 0.0 ? r : ri * pow(r, k - 2) * (k + log(r));
    return k * ri * pow(r, k - 2);
}
static double ga(double r, double eps) { return exp(-pow(r * eps, 2)); }
static double L_ga(double r, double eps, int k)
{
    double p[SERBF_REF_MAX_HV_ORDER + 1];
    double epr2 = pow(eps * r, 2);
    p[0] = 1;
    p[1] = 4 * epr2 - 4;
    for (int i = 1; i < k; i++)
        p[i + 1] = 4 * (epr2 - (2 * i + 1)) * p[i] - (16 * pow(i, 2)) * p[i - 1];
    return pow(eps, 2 * k) * p[k] * ga(r, eps);
}

int orc_lu_factor(int m, double *A, int *ipiv)
{
    int info = 0;
    for (int k = 0; k < m; k++) {
        int p = k;
        double amax = fabs(A[(size_t)k * m + k]);
        for (int i = k + 1; i < m; i++) {
            double a = fabs(A[(size_t)i * m + k]);
            if (a > amax) {
                amax = a;
                p = i;
            }
        }
        ipiv[k] = p;
        if (A[(size_t)p * m + k] == 0.0) {
            if (!info)
                info = k + 1;
            continue;
        }
        if (p != k)
            for (int j = 0; j < m; j++) {
                double t = A[(size_t)k * m + j];
                A[(size_t)k * m + j] = A[(size_t)p * m + j];
                A[(size_t)p * m + j] = t;
            }
        double piv = A[(size_t)k * m + k];
        for (int i = k + 1; i < m; i++) {
            double l = A[(size_t)i * m + k] / piv;
            A[(size_t)i * m + k] = l;
            if (l != 0.0)
                for (int j = k + 1; j < m; j++)
                    A[(size_t)i * m + j] -= l * A[(size_t)k * m + j];
        }
    }
    return info;
}

void orc_lu_solve(int m, int nrhs, const double *LU, const int *ipiv, double *B)
{
    for (int k = 0; k < m; k++) {
        int p = ipiv[k];
        if (p != k)
            for (int c = 0; c < nrhs; c++) {
                double t = B[(size_t)k * nrhs + c];
                B[(size_t)k * nrhs + c] = B[(size_t)p * nrhs + c];
                B[(size_t)p * nrhs + c] = t;
            }
    }
    for (int i = 0; i < m; i++)
        for (int j = 0; j < i; j++) {
            double l = LU[(size_t)i * m + j];
            for (int c = 0; c < nrhs; c++)
                B[(size_t)i * nrhs + c] -= l * B[(size_t)j * nrhs + c];
        }
    for (int i = m - 1; i >= 0; i--) {
        for (int j = i + 1; j < m; j++) {
            double u = LU[(size_t)i * m + j];
            for (int c = 0; c < nrhs; c++)
                B[(size_t)i * nrhs + c] -= u * B[(size_t)j * nrhs + c];
        }
        for (int c = 0; c < nrhs; c++)
            B[(size_t)i * nrhs + c] /= LU[(size_t)i * m + i];
    }
}

int orc_dgesv(int m, int nrhs, double *A, int *ipiv, double *B)
{
    int info = orc_lu_factor(m, A, ipiv);
    if (info)
        return info;
    orc_lu_solve(m, nrhs, A, ipiv, B);
    return 0;
}

/* One node of get_rbffd_DMs (rbffd_DMs.c:159-256).  X is 3 x n (row c = coordinate
 * c of the n stencil nodes, centre first).  Writes the rotated-frame weights and
 * the Gaussian Delta^k weights B2 (not yet divided by R).  Returns the info of the
 * PHS system; *info2 gets the info of the Gaussian system. */
int orc_ref_node_weights(int n, const double *X, const double H[9], int k_phs, int k_L,
                         double eps, double *hDxp, double *hDyp, double *Lw, int *info2)
{
    const int npoly = SERBF_REF_NPOLY(SERBF_REF_POLY_ORDER);
    const int size = n + npoly, size2 = n + 1;
    double *Xp = malloc(sizeof(double) * 3 * n);
    double *A = calloc((size_t)size * size, sizeof(double));
    double *A2 = calloc((size_t)size2 * size2, sizeof(double));
    double *B = calloc((size_t)size * 3, sizeof(double));
    double *B2 = calloc((size_t)size2, sizeof(double));
    int *ipiv = malloc(sizeof(int) * (size + size2));

    /* rbffd_DMs.c:168 Xp = H X */
    for (int r = 0; r < 3; r++)
        for (int s = 0; s < n; s++)
            Xp[r * n + s] = H[3 * r + 0] * X[0 * n + s] + H[3 * r + 1] * X[1 * n + s] +
                            H[3 * r + 2] * X[2 * n + s];
    const double *xp = Xp, *yp = Xp + n, *zp = Xp + 2 * n;

    /* rbffd_DMs.c:175-200 */
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) {
            double dx = xp[j] - xp[i], dy = yp[j] - yp[i], dz = zp[j] - zp[i];
            double r = sqrt(dx * dx + dy * dy + dz * dz);
            A[(size_t)i * size + j] = phs(r, SERBF_REF_PHS_ORDER);
            A2[(size_t)i * size2 + j] = ga(r, eps);
        }
        double dx = xp[0] - xp[i], dy = yp[0] - yp[i], dz = zp[0] - zp[i];
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        B[3 * i] = d1_phs(r, dx, k_phs);
        B[3 * i + 1] = d1_phs(r, dy, k_phs);
        B[3 * i + 2] = 0.0;
        B2[i] = L_ga(r, eps, k_L);
    }
    /* rbffd_DMs.c:202-215: monomials x'^(p-i) y'^i, p = 0..5 */
    int off = n;
    for (int p = 0; p <= SERBF_REF_POLY_ORDER; p++) {
        for (int i = 0; i <= p; i++) {
            for (int j = 0; j < n; j++) {
                double a = pow(xp[j], p - i) * pow(yp[j], i);
                A[(size_t)(i + off) * size + j] = a;
                A[(size_t)j * size + i + off] = a;
            }
            B[3 * (i + off)] = (p == 1 && i == 0) ? 1.0 : 0.0;
            B[3 * (i + off) + 1] = (p == 1 && i == 1) ? 1.0 : 0.0;
            B[3 * (i + off) + 2] = 0.0;
        }
        off += p + 1;
    }
    /* rbffd_DMs.c:223-228: Gaussian + constant */
    for (int i = 0; i < n; i++) {
        A2[(size_t)n * size2 + i] = 1.0;
        A2[(size_t)i * size2 + n] = 1.0;
    }
    A2[(size_t)size2 * size2 - 1] = 0.0;
    B2[n] = 0.0;

    int info1 = orc_dgesv(size, 3, A, ipiv, B);
    *info2 = orc_dgesv(size2, 1, A2, ipiv + size, B2);
    for (int s = 0; s < n; s++) {
        hDxp[s] = B[3 * s];
        hDyp[s] = B[3 * s + 1];
        Lw[s] = B2[s];
    }
    free(Xp); free(A); free(A2); free(B); free(B2); free(ipiv);
    return info1;
}

/* 2D Laplacian applied k times to the radial function r^K:
 * Delta r^m = m^2 r^(m-2), so Delta^k r^K = prod_{j<k} (K-2j)^2 r^(K-2k). */
static double lap_k_phs_coef(int K, int k)
{
    double c = 1.0;
    for (int j = 0; j < k; j++)
        c *= (double)(K - 2 * j) * (double)(K - 2 * j);
    return c;
}

/* Config-mode node weights.  Stencil rotated by H (pole fixed), shifted so the
 * centre is the origin and scaled by the stencil radius s; PHS r^k + monomials
 * x^(p-i) y^i (p <= d).  Gradient weights are divided by s, Delta^k weights by s^2k. */
static int cfg_node_weights(const serbf_config *cfg, int n, const double *X, const double H[9],
                            double *wx, double *wy, double *wl, double *wl1, double *wl2,
                            int eps_nmax, int *info2)
{
    const int d = cfg->poly_deg, M = (d + 1) * (d + 2) / 2, m = n + M;
    double *xs = malloc(sizeof(double) * 3 * n);
    double *A = malloc(sizeof(double) * (size_t)m * m);
    double *B = malloc(sizeof(double) * (size_t)m * 4);
    double *A2 = malloc(sizeof(double) * (size_t)m * m);
    double *B2 = malloc(sizeof(double) * (size_t)m * 3);
    int *ipiv = malloc(sizeof(int) * m);
    double c0[3];
    for (int r = 0; r < 3; r++)
        c0[r] = H[3 * r] * X[0] + H[3 * r + 1] * X[n] + H[3 * r + 2] * X[2 * n];
    double s = 0.0;
    for (int j = 0; j < n; j++) {
        double q[3];
        for (int r = 0; r < 3; r++)
            q[r] = H[3 * r] * X[j] + H[3 * r + 1] * X[n + j] + H[3 * r + 2] * X[2 * n + j] - c0[r];
        xs[j] = q[0]; xs[n + j] = q[1]; xs[2 * n + j] = q[2];
        double rr = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        if (rr > s)
            s = rr;
    }
    for (int j = 0; j < 3 * n; j++)
        xs[j] /= s;
    const double *px = xs, *py = xs + n, *pz = xs + 2 * n;
    const int k = cfg->phs_order, K = cfg->hv_phs_order, kk = cfg->hv_order;
    /* hv_phs_order == 0: Gaussian + constant Delta^k system (the reference's form,
     * rbffd_DMs.c:188,223-228, L_garbf rbffd_DMs.c:314-328) with the shape parameter
     * hv_eps of the unit sphere, i.e. eps * s on the scaled stencil */
    const int gauss = kk > 0 && K == 0;
    /* eps * s, and with per-node stencil sizes (n < cfg->n here only through orc_weights_var,
     * which passes the row stride in eps_n) eps grows as sqrt(n_max / n) so that eps times the
     * stencil radius is the same for every stencil size */
    const double geps = cfg->hv_eps * s * (eps_nmax > n ? sqrt((double)eps_nmax / n) : 1.0);
    const double cL = kk > 0 ? lap_k_phs_coef(K, kk) : 0.0;
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) {
            double dx = px[j] - px[i], dy = py[j] - py[i], dz = pz[j] - pz[i];
            double r = sqrt(dx * dx + dy * dy + dz * dz);
            A[(size_t)i * m + j] = pow(r, k);
            A2[(size_t)i * m + j] = kk > 0 ? (gauss ? ga(r, geps) : pow(r, K)) : 0.0;
        }
        double dx = -px[i], dy = -py[i], dz = -pz[i];
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        /* d/dx', d/dy', Delta, Delta^2 of r^k at the centre */
        B[4 * i] = k * dx * pow(r, k - 2);
        B[4 * i + 1] = k * dy * pow(r, k - 2);
        B[4 * i + 2] = lap_k_phs_coef(k, 1) * pow(r, k - 2);
        B[4 * i + 3] = lap_k_phs_coef(k, 2) * pow(r, k - 4);
        /* right-hand sides Delta^kk, Delta, Delta^2 of r^K at the centre */
        if (gauss) {
            B2[i] = L_ga(r, geps, kk);
        } else {
            B2[i] = kk > 0 ? cL * pow(r, K - 2 * kk) : 0.0;
        }
    }
    int off = n;
    for (int p = 0; p <= d; p++) {
        for (int i = 0; i <= p; i++) {
            for (int j = 0; j < n; j++) {
                double a = pow(px[j], p - i) * pow(py[j], i);
                A[(size_t)(off + i) * m + j] = a;
                A[(size_t)j * m + off + i] = a;
                A2[(size_t)(off + i) * m + j] = a;
                A2[(size_t)j * m + off + i] = a;
            }
            B[4 * (off + i)] = (p == 1 && i == 0) ? 1.0 : 0.0;
            B[4 * (off + i) + 1] = (p == 1 && i == 1) ? 1.0 : 0.0;
            /* Delta, Delta^2 of x^(p-i) y^i at the origin: 2 for x^2, y^2; 24, 8, 24 for
             * x^4, x^2 y^2, y^4; Delta^kk (kk >= 3) of degree < 2kk vanishes there */
            B[4 * (off + i) + 2] = (p == 2 && (i == 0 || i == 2)) ? 2.0 : 0.0;
            B[4 * (off + i) + 3] = p == 4 ? ((i == 0 || i == 4) ? 24.0 : (i == 2 ? 8.0 : 0.0)) : 0.0;
            B2[off + i] = 0.0;
        }
        off += p + 1;
    }
    for (int i = n; i < m; i++)
        for (int j = n; j < m; j++) {
            A[(size_t)i * m + j] = 0.0;
            A2[(size_t)i * m + j] = 0.0;
        }
    int info1 = orc_dgesv(m, 4, A, ipiv, B);
    *info2 = 0;
    if (gauss) {
        /* Gaussian + constant: rows/columns 0..n, the constant row's right-hand sides 0 */
        const int m2 = n + 1;
        double *G = malloc(sizeof(double) * (size_t)m2 * m2);
        for (int i = 0; i < n; i++) {
            for (int j = 0; j < n; j++)
                G[(size_t)i * m2 + j] = A2[(size_t)i * m + j];
            G[(size_t)i * m2 + n] = 1.0;
            G[(size_t)n * m2 + i] = 1.0;
        }
        G[(size_t)n * m2 + n] = 0.0;
        B2[n] = 0.0;
        *info2 = orc_dgesv(m2, 1, G, ipiv, B2);
        free(G);
    } else if (kk > 0)
        *info2 = orc_dgesv(m, 1, A2, ipiv, B2);
    double s2k = pow(s, 2 * kk);
    for (int j = 0; j < n; j++) {
        wx[j] = B[4 * j] / s;
        wy[j] = B[4 * j + 1] / s;
        wl[j] = kk > 0 ? B2[j] / s2k : 0.0;
        if (wl1) wl1[j] = B[4 * j + 2] / (s * s);
        if (wl2) wl2[j] = B[4 * j + 3] / (s * s * s * s);
    }
    free(xs); free(A); free(B); free(A2); free(B2); free(ipiv);
    return info1;
}

int orc_weights(const serbf_config *cfg, int N, const double *x, const double *y,
                const double *z, const int *idx, double *Dx, double *Dy, double *L,
                int *status)
{
    return orc_weights_lap(cfg, N, x, y, z, idx, Dx, Dy, L, NULL, NULL, status);
}

int orc_weights_lap(const serbf_config *cfg, int N, const double *x, const double *y,
                    const double *z, const int *idx, double *Dx, double *Dy, double *L,
                    double *L1, double *L2, int *status)
{
    const int n = cfg->n;
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 32) reduction(+ : bad)
    for (int i = 0; i < N; i++) {
        double X[3 * 512];
        double H[9];
        const int *st = idx + (size_t)i * n;
        for (int s = 0; s < n; s++) {
            X[s] = x[st[s]];
            X[n + s] = y[st[s]];
            X[2 * n + s] = z[st[s]];
        }
        int info1, info2;
        if (cfg->mode == SERBF_MODE_REFERENCE) {
            orc_householder(x[i], y[i], z[i], 0, H);
            info1 = orc_ref_node_weights(n, X, H, cfg->phs_order, cfg->hv_order, cfg->hv_eps,
                                         Dx + (size_t)i * n, Dy + (size_t)i * n,
                                         L + (size_t)i * n, &info2);
        } else {
            orc_householder(x[i], y[i], z[i], 1, H);
            info1 = cfg_node_weights(cfg, n, X, H, Dx + (size_t)i * n, Dy + (size_t)i * n,
                                     L + (size_t)i * n, L1 ? L1 + (size_t)i * n : NULL,
                                     L2 ? L2 + (size_t)i * n : NULL, n, &info2);
        }
        status[i] = info1 ? 1 : (info2 ? 2 : 0);
        bad += status[i] != 0;
    }
    return bad;
}

/* Config mode with a stencil size nn[i] <= cfg->n and polynomial degree deg[i] per node; the
 * stencil and weight rows keep the stride cfg->n, entries past nn[i] get zero weights. */
int orc_weights_var(const serbf_config *cfg, int N, const double *x, const double *y,
                    const double *z, const int *idx, const int *nn, const int *deg, double *Dx,
                    double *Dy, double *L, int *status)
{
    const int stride = cfg->n;
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 32) reduction(+ : bad)
    for (int i = 0; i < N; i++) {
        serbf_config c = *cfg;
        const int n = nn[i];
        c.n = n;
        c.poly_deg = deg[i];
        double X[3 * 512];
        double H[9];
        const int *st = idx + (size_t)i * stride;
        for (int s = 0; s < n; s++) {
            X[s] = x[st[s]];
            X[n + s] = y[st[s]];
            X[2 * n + s] = z[st[s]];
        }
        orc_householder(x[i], y[i], z[i], 1, H);
        double *dx = Dx + (size_t)i * stride, *dy = Dy + (size_t)i * stride, *dl = L + (size_t)i * stride;
        int info2 = 0;
        const int info1 = cfg_node_weights(&c, n, X, H, dx, dy, dl, NULL, NULL, stride, &info2);
        for (int s = n; s < stride; s++)
            dx[s] = dy[s] = dl[s] = 0.0;
        status[i] = info1 ? 1 : (info2 ? 2 : 0);
        bad += status[i] != 0;
    }
    return bad;
}

/* Vertical operator.  Reference mode: 7-point 6th-order FD1 / dh on levels v-3..v+3
 * with zero values outside the column (config.h:21-22, rbffd_DMs.c:146-154,412-427).
 * Config mode: 1D PHS |z|^kv + degree-dv polynomials on the nv levels nearest to v
 * (centred in the interior, one-sided at the ends), on coordinates in units of dh. */
int orc_vertical(const serbf_config *cfg, int *start, double *W)
{
    const int Nv = cfg->Nv;
    const double dh = cfg->htop / Nv;
    if (cfg->mode == SERBF_MODE_REFERENCE) {
        static const double fd1[7] = {-1.0 / 60.0, 3.0 / 20.0, -3.0 / 4.0, 0.0,
                                      3.0 / 4.0, -3.0 / 20.0, 1.0 / 60.0};
        for (int v = 0; v < Nv; v++) {
            start[v] = v - 3;
            for (int j = 0; j < 7; j++)
                W[v * 7 + j] = fd1[j] / dh;
        }
        return 7;
    }
    const int nv = cfg->nv, kv = cfg->v_phs_order, dv = cfg->v_poly_deg, m = nv + dv + 1;
    if (nv > Nv || nv < 1)
        return -1;
    double A[32 * 32], B[32];
    int ipiv[32];
    for (int v = 0; v < Nv; v++) {
        int s0 = v - nv / 2;
        if (s0 < 0)
            s0 = 0;
        if (s0 > Nv - nv)
            s0 = Nv - nv;
        start[v] = s0;
        double zc[16];
        for (int j = 0; j < nv; j++)
            zc[j] = (double)(s0 + j - v);
        memset(A, 0, sizeof(A));
        for (int i = 0; i < nv; i++) {
            for (int j = 0; j < nv; j++)
                A[i * m + j] = pow(fabs(zc[i] - zc[j]), kv);
            for (int a = 0; a <= dv; a++) {
                double p = pow(zc[i], a);
                A[i * m + nv + a] = p;
                A[(nv + a) * m + i] = p;
            }
            B[i] = kv * (0.0 - zc[i]) * pow(fabs(zc[i]), kv - 2);
        }
        for (int a = 0; a <= dv; a++)
            B[nv + a] = (a == 1) ? 1.0 : 0.0;
        if (orc_dgesv(m, 1, A, ipiv, B))
            return -1;
        for (int j = 0; j < nv; j++)
            W[v * nv + j] = B[j] / dh;
    }
    return nv;
}
