/*
 * Oracle flows, initial conditions, right-hand side, RK4 and norms (test
 * infrastructure only).
 *
 * RHS per node p and level v (eval_partSVDotGradSV, rbffd_DMs.c:370-442):
 *   gx = r_inv[v] * sum_s Dx[p,s] q[idx[p,s], v]       (sum in stencil order)
 *   gy = r_inv[v] * sum_s Dy[p,s] q[idx[p,s], v]
 *   gz = sum_j W[v,j] q[p, start[v]+j]                  (zero outside the column)
 *   F  = -1.0 * ((Up_x gx + Up_y gy) + Up_z gz) + c * sum_s L[p,s] q[idx[p,s], v]
 * with c = hv_coef (or a per-point coefficient), Up the velocity in the node's
 * Householder frame.  RK4 follows get_solution (wrappers.c:54-104) expression by
 * expression: F = F1; F = F + 2F2; F = F + 2F3; F = F + 1.0F4; q = q + dt/6 F.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* OpenMP team size of the oracle's parallel loops (the CPU-baseline timing sets and
 * reports it) */
void orc_set_threads(int n)
{
    if (n > 0)
        omp_set_num_threads(n);
}
int orc_max_threads(void) { return omp_get_max_threads(); }

static double rho_of(double h)
{
    const double rho0 = SERBF_P0 / (SERBF_RD * SERBF_T0);
    return rho0 * exp(-(h * SERBF_GRAVITY) / (SERBF_RD * SERBF_T0));
}

static void lonlat(double x, double y, double z, double *lam, double *phi)
{
    double l = atan2(y, x);
    if (l < 0.0)
        l += 2.0 * SERBF_PI;
    *lam = l;
    *phi = atan2(z, sqrt(x * x + y * y));
}

/* (u_lambda, u_phi, u_r) of the flow at longitude lam, latitude phi, height h. */
static void vel_sph(const serbf_config *cfg, double lam, double phi, double h, double t,
                    double *ul, double *up, double *ur)
{
    if (cfg->flow == SERBF_FLOW_TC2) {
        /* test_cases.c:131-145,186-228 (tau = T, K = k_def) */
        const double rho0 = SERBF_P0 / (SERBF_RD * SERBF_T0);
        const double K = cfg->k_def, w0 = cfg->w0, htop = cfg->htop, R = cfg->R;
        double wt = cos(SERBF_PI * (t / cfg->T));
        double u_r = (w0 * rho0) / (K * rho_of(h)) *
                     (((-2) * sin(K * phi) * sin(phi)) + (K * cos(K * phi) * cos(phi))) *
                     sin(SERBF_PI * h / htop);
        double u_phi = -((R * w0 * SERBF_PI * rho0) / (K * htop * rho_of(h))) * cos(phi) *
                       sin(K * phi) * cos(SERBF_PI * h / htop);
        *ur = u_r * wt;
        *ul = cfg->u0 * cos(phi);
        *up = u_phi * wt;
    } else if (cfg->flow == SERBF_FLOW_SOLID_BODY) {
        /* u = u0 (cos a cos phi + sin a cos lam sin phi), v = -u0 sin a sin lam */
        double a = cfg->alpha;
        *ul = cfg->u0 * (cos(a) * cos(phi) + sin(a) * cos(lam) * sin(phi));
        *up = -cfg->u0 * sin(a) * sin(lam);
        *ur = 0.0;
    } else {
        /* solid-body rotation + reversing swirl in the rotating frame + reversing w */
        double T = cfg->T;
        double lp = lam - 2.0 * SERBF_PI * t / T;
        double ct = cos(SERBF_PI * t / T);
        double sl = sin(lp);
        *ul = cfg->k_def * sl * sl * sin(2.0 * phi) * ct + cfg->u0 * cos(phi);
        *up = cfg->k_def * sin(2.0 * lp) * cos(phi) * ct;
        *ur = cfg->w0 * sin(SERBF_PI * h / cfg->htop) * ct;
    }
}

void orc_velocity_cart(const serbf_config *cfg, double x, double y, double z, double h,
                       double t, double U[3])
{
    double lam, phi, ul, up, ur;
    lonlat(x, y, z, &lam, &phi);
    vel_sph(cfg, lam, phi, h, t, &ul, &up, &ur);
    /* test_cases.c:143-145 */
    U[0] = cos(phi) * cos(lam) * ur - sin(lam) * ul - sin(phi) * cos(lam) * up;
    U[1] = cos(phi) * sin(lam) * ur + cos(lam) * ul - sin(phi) * sin(lam) * up;
    U[2] = sin(phi) * ur + cos(phi) * up;
}

void orc_velocity(const serbf_config *cfg, double x, double y, double z, double h, double t,
                  double Up[3])
{
    double U[3], H[9];
    int ref = cfg->mode == SERBF_MODE_REFERENCE;
    orc_velocity_cart(cfg, x, y, z, h, t, U);
    orc_householder(x, y, z, !ref, H);
    for (int r = 0; r < 3; r++)
        Up[r] = H[3 * r] * U[0] + H[3 * r + 1] * U[1] + H[3 * r + 2] * U[2];
    if (!ref) {
        double lam, phi, ul, up, ur;
        lonlat(x, y, z, &lam, &phi);
        vel_sph(cfg, lam, phi, h, t, &ul, &up, &ur);
        Up[2] = ur;
    }
}

/* chord distance on the unit sphere between (x,y,z) and (lam_c, phi_c) */
static double chord_to(double x, double y, double z, double lc, double pc)
{
    double cx = cos(pc) * cos(lc), cy = cos(pc) * sin(lc), cz = sin(pc);
    double dx = x - cx, dy = y - cy, dz = z - cz;
    return sqrt(dx * dx + dy * dy + dz * dz);
}
/* great-circle angle from the chord */
static double angle_of_chord(double c) { return 2.0 * asin(fmin(1.0, 0.5 * c)); }

static unsigned long long splitmix(unsigned long long *s)
{
    unsigned long long z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double unif(unsigned long long *s) { return (splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

void orc_initial(int kind, const serbf_config *cfg, int N, const double *x, const double *y,
                 const double *z, unsigned long long seed, double *q)
{
    const int Nv = cfg->Nv;
    const double dh = cfg->htop / Nv;
    double lc[2], pc[2];
    unsigned long long s = seed;
    for (int k = 0; k < 2; k++) {
        lc[k] = 2.0 * SERBF_PI * unif(&s);
        pc[k] = asin(2.0 * unif(&s) - 1.0) * (2.0 / 3.0);
    }
    for (int p = 0; p < N; p++) {
        double lam, phi;
        lonlat(x[p], y[p], z[p], &lam, &phi);
        for (int v = 0; v < Nv; v++) {
            double h = dh * (v + 0.5), val = 0.0;
            switch (kind) {
            case ORC_IC_TC2: {  /* test_cases.c:171-180 */
                double h1 = 2000.0, h2 = 5000.0, h0 = (h1 + h2) / 2;
                val = h < h1 || h > h2 ? 0.0 : (1.0 + cos((2.0 * SERBF_PI) * (h - h0) / (h2 - h1))) / 2.0;
                break;
            }
            case ORC_IC_BELL: {
                double a = angle_of_chord(chord_to(x[p], y[p], z[p], 1.5 * SERBF_PI, 0.0));
                double rb = 1.0 / 3.0;
                double bell = a < rb ? 0.5 * (1.0 + cos(SERBF_PI * a / rb)) : 0.0;
                double g = (h - 5000.0) / 1000.0;
                val = bell * exp(-0.5 * g * g);
                break;
            }
            case ORC_IC_HILLS: {
                double g = (h - 0.5 * cfg->htop) / (0.2 * cfg->htop);
                double sum = 0.0;
                for (int k = 0; k < 2; k++) {
                    double c = chord_to(x[p], y[p], z[p], lc[k], pc[k]) / 0.2;
                    sum += exp(-c * c);
                }
                val = sum * exp(-g * g);
                break;
            }
            case ORC_IC_SLOTTED: {
                double r = 0.5, lcs[2] = {5.0 * SERBF_PI / 6.0, 7.0 * SERBF_PI / 6.0};
                val = 0.1;
                /* two cylinders of radius r/2 rad with a slot of half-width r/6
                 * cut from latitude -5r/12 northwards */
                for (int k = 0; k < 2; k++) {
                    double a = angle_of_chord(chord_to(x[p], y[p], z[p], lcs[k], 0.0));
                    if (a <= r && (fabs(lam - lcs[k]) >= r / 6.0 || phi < -5.0 * r / 12.0))
                        val = 1.0;
                }
                break;
            }
            case ORC_IC_BELLS2: {
                double r = 0.5, lcs[2] = {5.0 * SERBF_PI / 6.0, 7.0 * SERBF_PI / 6.0};
                val = 0.1;
                for (int k = 0; k < 2; k++) {
                    double a = angle_of_chord(chord_to(x[p], y[p], z[p], lcs[k], 0.0));
                    if (a < r)
                        val += 0.9 * 0.5 * (1.0 + cos(SERBF_PI * a / r));
                }
                break;
            }
            default:
                val = 1.0;
            }
            q[(size_t)p * Nv + v] = val;
        }
    }
}

void orc_rhs(const serbf_config *cfg, int N, const double *x, const double *y, const double *z,
             const int *idx, const double *Dx, const double *Dy, const double *L,
             const int *vstart, const double *vW, const double *gamma, const double *q,
             double t, double *F)
{
    const int Nv = cfg->Nv, n = cfg->n, ntr = cfg->ntracers > 0 ? cfg->ntracers : 1;
    const int nv = cfg->mode == SERBF_MODE_REFERENCE ? 7 : cfg->nv;
    const double dh = cfg->htop / Nv;
    const int use_hv = cfg->hv_order > 0 && (cfg->hv_coef != 0.0 || gamma);
#pragma omp parallel for schedule(static)
    for (int p = 0; p < N; p++) {
        const int *st = idx + (size_t)p * n;
        for (int tr = 0; tr < ntr; tr++) {
            const size_t col = ((size_t)p * ntr + tr) * Nv;
            for (int v = 0; v < Nv; v++) {
                double h = dh * (v + 0.5);
                double r_inv = 1 / (cfg->R + h);
                double ax = 0.0, ay = 0.0, al = 0.0;
                for (int s = 0; s < n; s++) {
                    double b = q[((size_t)st[s] * ntr + tr) * Nv + v];
                    ax += Dx[(size_t)p * n + s] * b;
                    ay += Dy[(size_t)p * n + s] * b;
                    if (use_hv)
                        al += L[(size_t)p * n + s] * b;
                }
                ax *= r_inv;
                ay *= r_inv;
                double az = 0.0;
                for (int j = 0; j < nv; j++) {
                    int lv = vstart[v] + j;
                    double b = (lv >= 0 && lv < Nv) ? q[col + lv] : 0.0;
                    az += vW[v * nv + j] * b;
                }
                double Up[3];
                orc_velocity(cfg, x[p], y[p], z[p], h, t, Up);
                double f = -1.0 * ((Up[0] * ax + Up[1] * ay) + Up[2] * az);
                if (use_hv)
                    f += (gamma ? gamma[col + v] : cfg->hv_coef) * al;
                F[col + v] = f;
            }
        }
    }
}

void orc_rk4(const serbf_config *cfg, int N, const double *x, const double *y, const double *z,
             const int *idx, const double *Dx, const double *Dy, const double *L,
             const int *vstart, const double *vW, int nsteps, double dt, double t0, double *q)
{
    const size_t P = (size_t)N * (cfg->ntracers > 0 ? cfg->ntracers : 1) * cfg->Nv;
    double *F = malloc(P * sizeof(double)), *Fk = malloc(P * sizeof(double)),
           *qk = malloc(P * sizeof(double));
    for (int step = 0; step < nsteps; step++) {
        double t = t0 + step * dt;
        orc_rhs(cfg, N, x, y, z, idx, Dx, Dy, L, vstart, vW, NULL, q, t, F);
        for (size_t i = 0; i < P; i++) qk[i] = q[i] + (dt / 2.0) * F[i];
        orc_rhs(cfg, N, x, y, z, idx, Dx, Dy, L, vstart, vW, NULL, qk, t + dt / 2.0, Fk);
        for (size_t i = 0; i < P; i++) F[i] = F[i] + 2.0 * Fk[i];
        for (size_t i = 0; i < P; i++) qk[i] = q[i] + (dt / 2.0) * Fk[i];
        orc_rhs(cfg, N, x, y, z, idx, Dx, Dy, L, vstart, vW, NULL, qk, t + dt / 2.0, Fk);
        for (size_t i = 0; i < P; i++) F[i] = F[i] + 2.0 * Fk[i];
        for (size_t i = 0; i < P; i++) qk[i] = q[i] + dt * Fk[i];
        orc_rhs(cfg, N, x, y, z, idx, Dx, Dy, L, vstart, vW, NULL, qk, t + dt, Fk);
        for (size_t i = 0; i < P; i++) F[i] = F[i] + 1.0 * Fk[i];
        for (size_t i = 0; i < P; i++) q[i] = q[i] + (dt / 6.0) * F[i];
    }
    free(F); free(Fk); free(qk);
}

/* verify_solution (wrappers.c:133-171) on flat arrays; all tracers pooled */
void orc_norms(const serbf_config *cfg, int N, const double *q_ref, const double *q,
               double out[7])
{
    const int Nv = cfg->Nv, ntr = cfg->ntracers > 0 ? cfg->ntracers : 1;
    const double dh = cfg->htop / Nv;
    double e[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 1.0};
    e[8] = -INFINITY;
    e[9] = INFINITY;
    for (int pt = 0; pt < N * ntr; pt++)
        for (int v = 0; v < Nv; v++) {
            double h = dh * (v + 0.5), r = cfg->R + h;
            double qr = q_ref[(size_t)pt * Nv + v], qa = q[(size_t)pt * Nv + v];
            double diff = fabs(qr - qa);
            e[0] += diff;
            e[1] += fabs(qr);
            e[2] += diff * diff;
            e[3] += qr * qr;
            double V = (4 * SERBF_PI * r * r * dh) / N;
            e[4] += rho_of(h) * V * qr;
            e[5] += rho_of(h) * V * qa;
            e[6] = e[6] > diff ? e[6] : diff;
            e[7] = e[7] > fabs(qa) ? e[7] : fabs(qa);
            e[8] = e[8] < qa ? qa : e[8];
            e[9] = e[9] < qa ? e[9] : qa;
        }
    out[0] = e[0] / e[1];
    out[1] = sqrt(e[2] / e[3]);
    out[2] = e[6] / e[7];
    out[3] = e[4];
    out[4] = e[5];
    out[5] = e[9];
    out[6] = e[8];
}
