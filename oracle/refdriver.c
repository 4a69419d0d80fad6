/*
 * Oracle: struct-level restatement of the reference solver's hot-path functions
 * on the reference structs (test infrastructure only).  Each function follows
 * the reference function named in its comment loop by loop, with MKL calls
 * replaced by the explicit sums they compute and MPI reduced to one rank.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* nodesets.c:113-153 (one rank: the slice is the whole node set) */
void orc_ref_get_stencils(nodeset_struct *ns, int n)
{
    ns->n = n;
    int N = ns->Nnodes;
    int *idx = malloc(sizeof(int) * (size_t)N * n);
    double *D = malloc(sizeof(double) * (size_t)N * n);
    orc_knn_brute(N, ns->x, ns->y, ns->z, n, idx, D);
    ns->idx = idx;
    ns->D = D;
}

/* layers.c:15-57 */
void orc_ref_get_layers(layers_struct *L, double htop, int Nv, double R)
{
    int Nvt = Nv + 2 * SERBF_REF_GHOST;
    int pNv = SERBF_ALIGNUP(Nv, SERBF_REF_CACHE_ALIGN);
    int pNvt = SERBF_ALIGNUP(Nvt, SERBF_REF_CACHE_ALIGN);
    double dh = htop / Nv;
    double *h = malloc(sizeof(double) * pNv), *r = malloc(sizeof(double) * pNv),
           *ri = malloc(sizeof(double) * pNv);
    for (int i = 0; i < pNv; i++) {
        double height = dh * (i + .5);
        h[i] = height;
        r[i] = R + height;
        ri[i] = 1 / (R + height);
    }
    L->Nv = Nv;
    L->Ng = SERBF_REF_GHOST;
    L->Nvt = Nvt;
    L->padded_Nv = pNv;
    L->padded_Nvt = pNvt;
    L->dh = dh;
    L->htop = htop;
    L->h = h;
    L->r = r;
    L->r_inv = ri;
}

/* rbffd_DMs.c:55-96 */
void orc_ref_init_HH(patch_struct *p)
{
    int Np = p->rbffd_DMs->part_Nnodes;
    double *H = calloc((size_t)Np * 9, sizeof(double));
    for (int i = 0; i < Np; i++) {
        int pid = p->part_pids[i];
        orc_householder(p->nodeset->x[pid], p->nodeset->y[pid], p->nodeset->z[pid], 0, H + 9 * i);
    }
    p->rbffd_DMs->H = H;
}

/* rbffd_DMs.c:98-282 */
void orc_ref_get_rbffd_DMs(patch_struct *p, int k_phs, int k_L, double eps, double R)
{
    rbffd_DMs_struct *dm = p->rbffd_DMs;
    int Np = dm->part_Nnodes, n = dm->n;
    const int F7 = SERBF_REF_FD1_SIZE;
    static const double fd1[7] = {-1.0 / 60.0, 3.0 / 20.0, -3.0 / 4.0, 0.0,
                                  3.0 / 4.0, -3.0 / 20.0, 1.0 / 60.0};
    double vDzp[7];
    for (int i = 0; i < F7; i++)
        vDzp[i] = fd1[i] / p->layers->dh;
    double *hDx = malloc(sizeof(double) * Np * n), *hDy = malloc(sizeof(double) * Np * n),
           *hDz = malloc(sizeof(double) * Np * n), *L = malloc(sizeof(double) * Np * n),
           *hDxp = malloc(sizeof(double) * Np * n), *hDyp = malloc(sizeof(double) * Np * n),
           *vDx = malloc(sizeof(double) * Np * F7), *vDy = malloc(sizeof(double) * Np * F7),
           *vDz = malloc(sizeof(double) * Np * F7);
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < Np; i++) {
        double *X = malloc(sizeof(double) * 3 * n);
        for (int s = 0; s < n; s++) {
            int pid = dm->idx[i * n + s];
            X[s] = p->nodeset->x[pid];
            X[n + s] = p->nodeset->y[pid];
            X[2 * n + s] = p->nodeset->z[pid];
        }
        const double *H = dm->H + 9 * i;
        int info2;
        orc_ref_node_weights(n, X, H, k_phs, k_L, eps, hDxp + i * n, hDyp + i * n, L + i * n,
                             &info2);
        for (int s = 0; s < n; s++) {
            /* W = H B^T (rbffd_DMs.c:241), B's third column is zero */
            double b0 = hDxp[i * n + s], b1 = hDyp[i * n + s];
            hDx[i * n + s] = H[0] * b0 + H[1] * b1;
            hDy[i * n + s] = H[3] * b0 + H[4] * b1;
            hDz[i * n + s] = H[6] * b0 + H[7] * b1;
            L[i * n + s] /= R;
        }
        for (int v = 0; v < F7; v++) {
            vDx[i * F7 + v] = H[2] * vDzp[v];
            vDy[i * F7 + v] = H[5] * vDzp[v];
            vDz[i * F7 + v] = H[8] * vDzp[v];
        }
        free(X);
    }
    dm->hDx = hDx; dm->hDy = hDy; dm->hDz = hDz;
    dm->vDx = vDx; dm->vDy = vDy; dm->vDz = vDz;
    dm->L = L; dm->hDxp = hDxp; dm->hDyp = hDyp;
    memcpy(dm->vDzp, vDzp, sizeof(vDzp));
}

/* rbffd_DMs.c:331-443 */
void orc_ref_eval_partSVDotGradSV(patch_struct *p, double alpha, part_SV_struct *SA,
                                  patch_SV_struct *SB, part_SV_struct *SC)
{
    layers_struct *ly = p->layers;
    rbffd_DMs_struct *dm = p->rbffd_DMs;
    int Np = p->part_Nnodes, Nv = ly->Nv, Ng = ly->Ng, pNv = ly->padded_Nv, pNvt = ly->padded_Nvt;
    int n = dm->n;
    const double *A = SA->data, *B = SB->data;
    double *C = SC->data, *g = p->SV_3D_temp->data;
    for (int i = 0; i < 3 * pNv * Np; i++)
        g[i] = 0.0;
    for (int i = 0; i < Np; i++) {
        int pid = p->part_pids[i];
        for (int s = 0; s < n; s++) {
            int nb = dm->idx[i * n + s];
            double wx = dm->hDxp[i * n + s], wy = dm->hDyp[i * n + s];
            for (int v = 0; v < Nv; v++) {
                double b = B[nb * pNvt + v];
                g[(3 * i + 0) * pNv + v] += wx * b;
                g[(3 * i + 1) * pNv + v] += wy * b;
            }
        }
        for (int v = 0; v < Nv; v++) {
            g[(3 * i + 0) * pNv + v] *= ly->r_inv[v];
            g[(3 * i + 1) * pNv + v] *= ly->r_inv[v];
        }
        for (int s = 0; s < SERBF_REF_FD1_SIZE; s++) {
            double w = dm->vDzp[s];
            for (int v = 0; v < Nv; v++)
                g[(3 * i + 2) * pNv + v] += w * B[pid * pNvt + v + (s - Ng)];
        }
    }
    for (int i = 0; i < Np; i++)
        for (int v = 0; v < Nv; v++) {
            double u = A[(3 * i + 0) * pNv + v] * g[(3 * i + 0) * pNv + v];
            double vv = A[(3 * i + 1) * pNv + v] * g[(3 * i + 1) * pNv + v];
            double w = A[(3 * i + 2) * pNv + v] * g[(3 * i + 2) * pNv + v];
            C[i * pNv + v] = alpha * (u + vv + w);
        }
}

/* rbffd_DMs.c:446-470 */
void orc_ref_eval_sumpatchSVpartSV(patch_struct *p, patch_SV_struct *SA, double alpha,
                                   part_SV_struct *SB, patch_SV_struct *SC)
{
    int Np = p->part_Nnodes, Nv = p->layers->Nv, pNv = p->layers->padded_Nv,
        pNvt = p->layers->padded_Nvt;
    for (int i = 0; i < Np; i++) {
        int pid = p->part_pids[i];
        for (int v = 0; v < Nv; v++)
            SC->data[pid * pNvt + v] = SA->data[pid * pNvt + v] + alpha * SB->data[i * pNv + v];
    }
}

/* rbffd_DMs.c:500-520 */
void orc_ref_eval_sumpartSVpartSV(patch_struct *p, part_SV_struct *SA, double alpha,
                                  part_SV_struct *SB, part_SV_struct *SC)
{
    int Np = p->part_Nnodes, Nv = p->layers->Nv, pNv = p->layers->padded_Nv;
    for (int i = 0; i < Np; i++)
        for (int v = 0; v < Nv; v++)
            SC->data[i * pNv + v] = SA->data[i * pNv + v] + alpha * SB->data[i * pNv + v];
}

/* rbffd_DMs.c:523-544: Up = H U over all padded_Nv columns */
void orc_ref_update_Up(patch_struct *p)
{
    int Np = p->part_Nnodes, pNv = p->layers->padded_Nv;
    const double *U = p->SV_U->data;
    double *Up = p->SV_Up->data;
    for (int i = 0; i < Np; i++) {
        const double *H = p->rbffd_DMs->H + 9 * i;
        for (int r = 0; r < 3; r++)
            for (int v = 0; v < pNv; v++)
                Up[(3 * i + r) * pNv + v] = H[3 * r + 0] * U[(3 * i + 0) * pNv + v] +
                                            H[3 * r + 1] * U[(3 * i + 1) * pNv + v] +
                                            H[3 * r + 2] * U[(3 * i + 2) * pNv + v];
    }
}

/* TC2 constants (test_cases.c:6-11) */
static const double TC2_TAU_ = 86400.0, TC2_K_ = 5.0, TC2_U0_ = 40.0, TC2_W0_ = 0.15,
                    TC2_H1_ = 2000.0, TC2_H2_ = 5000.0;

static double rho_h(const phys_constants_struct *pc, double h)
{
    return pc->rho0 * exp(-(h * pc->g) / (pc->Rd * pc->T0));
}

/* test_cases.c:38-100 */
void orc_ref_init_TC2(patch_struct *p, rt_config_struct *rt, const phys_constants_struct *pc)
{
    double dt = rt->timestep_length, t_end = TC2_TAU_;
    int nsteps = (int)ceil(t_end / dt);
    rt->nsteps = nsteps;
    rt->dt = t_end / nsteps;
    layers_struct *ly = p->layers;
    int Nv = ly->Nv, pNv = ly->padded_Nv, pNvt = ly->padded_Nvt;
    double h0 = (TC2_H1_ + TC2_H2_) / 2;
    for (int pid = 0; pid < p->Nnodes; pid++)
        for (int v = 0; v < Nv; v++) {
            double h = ly->h[v];
            double q = h < TC2_H1_ || h > TC2_H2_ ? 0.0
                       : (1.0 + cos((2.0 * SERBF_PI) * (h - h0) / (TC2_H2_ - TC2_H1_))) / 2.0;
            p->SV_q_init->data[pid * pNvt + v] = q;
            p->SV_q_t->data[pid * pNvt + v] = q;
        }
    for (int i = 0; i < p->part_Nnodes; i++) {
        int pid = p->part_pids[i];
        double phi = p->nodeset->phi[pid];
        for (int v = 0; v < Nv; v++) {
            double h = ly->h[v];
            double ur = (TC2_W0_ * pc->rho0) / (TC2_K_ * rho_h(pc, h)) *
                        (((-2) * sin(TC2_K_ * phi) * sin(phi)) + (TC2_K_ * cos(TC2_K_ * phi) * cos(phi))) *
                        sin(SERBF_PI * h / pc->htop);
            double ul = TC2_U0_ * cos(phi);
            double uphi = -((pc->R * TC2_W0_ * SERBF_PI * pc->rho0) / (TC2_K_ * pc->htop * rho_h(pc, h))) *
                          cos(phi) * sin(TC2_K_ * phi) * cos(SERBF_PI * h / pc->htop);
            p->SV_Usph->data[(3 * i + 0) * pNv + v] = ur;
            p->SV_Usph->data[(3 * i + 1) * pNv + v] = ul;
            p->SV_Usph->data[(3 * i + 2) * pNv + v] = uphi;
        }
    }
}

/* test_cases.c:106-148 */
void orc_ref_set_TC2_U_t(patch_struct *p, double t, const phys_constants_struct *pc)
{
    (void)pc;
    int Nv = p->layers->Nv, pNv = p->layers->padded_Nv;
    double wt = cos(SERBF_PI * (t / TC2_TAU_));
    const double *Us = p->SV_Usph->data;
    double *U = p->SV_U->data;
    for (int i = 0; i < p->part_Nnodes; i++) {
        int pid = p->part_pids[i];
        double ph = p->nodeset->phi[pid], la = p->nodeset->lambda[pid];
        for (int v = 0; v < Nv; v++) {
            double ur = Us[(3 * i + 0) * pNv + v] * wt;
            double ul = Us[(3 * i + 1) * pNv + v];
            double up = Us[(3 * i + 2) * pNv + v] * wt;
            U[(3 * i + 0) * pNv + v] = cos(ph) * cos(la) * ur - sin(la) * ul - sin(ph) * cos(la) * up;
            U[(3 * i + 1) * pNv + v] = cos(ph) * sin(la) * ur + cos(la) * ul - sin(ph) * sin(la) * up;
            U[(3 * i + 2) * pNv + v] = sin(ph) * ur + cos(ph) * up;
        }
    }
}

/* wrappers.c:39-106 */
void orc_ref_get_solution(patch_struct *p, const rt_config_struct *rt,
                          const phys_constants_struct *pc)
{
    int nsteps = rt->nsteps;
    double dt = rt->dt;
    orc_ref_set_TC2_U_t(p, 0.0, pc);
    orc_ref_update_Up(p);
    for (int ts = 0; ts < nsteps; ts++) {
        double t = ts * dt;
        orc_ref_eval_partSVDotGradSV(p, -1.0, p->SV_Up, p->SV_q_t, p->SV_F);
        orc_ref_set_TC2_U_t(p, t + (dt / 2.0), pc);
        orc_ref_update_Up(p);
        orc_ref_eval_sumpatchSVpartSV(p, p->SV_q_t, dt / 2.0, p->SV_F, p->SV_q_k);
        orc_ref_eval_partSVDotGradSV(p, -1.0, p->SV_Up, p->SV_q_k, p->SV_F_k);
        orc_ref_eval_sumpartSVpartSV(p, p->SV_F, 2.0, p->SV_F_k, p->SV_F);
        orc_ref_eval_sumpatchSVpartSV(p, p->SV_q_t, dt / 2.0, p->SV_F_k, p->SV_q_k);
        orc_ref_eval_partSVDotGradSV(p, -1.0, p->SV_Up, p->SV_q_k, p->SV_F_k);
        orc_ref_eval_sumpartSVpartSV(p, p->SV_F, 2.0, p->SV_F_k, p->SV_F);
        orc_ref_set_TC2_U_t(p, t + dt, pc);
        orc_ref_update_Up(p);
        orc_ref_eval_sumpatchSVpartSV(p, p->SV_q_t, dt, p->SV_F_k, p->SV_q_k);
        orc_ref_eval_partSVDotGradSV(p, -1.0, p->SV_Up, p->SV_q_k, p->SV_F_k);
        orc_ref_eval_sumpartSVpartSV(p, p->SV_F, 1.0, p->SV_F_k, p->SV_F);
        orc_ref_eval_sumpatchSVpartSV(p, p->SV_q_t, dt / 6.0, p->SV_F, p->SV_q_t);
    }
}

/* wrappers.c:108-176 (one rank), including the reference's initial values of the
 * running extrema (max starts at 0.0, min at 1.0) */
void orc_ref_verify(patch_struct *p, int global_N, const phys_constants_struct *pc,
                    double out[7])
{
    layers_struct *ly = p->layers;
    int Nv = ly->Nv, pNvt = ly->padded_Nvt;
    double e[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 1.0};
    for (int i = 0; i < p->part_Nnodes; i++) {
        int pid = p->part_pids[i];
        for (int v = 0; v < Nv; v++) {
            double qr = p->SV_q_init->data[pid * pNvt + v], qa = p->SV_q_t->data[pid * pNvt + v];
            double diff = fabs(qr - qa);
            e[0] += diff;
            e[1] += fabs(qr);
            e[2] += pow(diff, 2);
            e[3] += pow(qr, 2);
            double rho = rho_h(pc, ly->h[v]);
            double V = (4 * SERBF_PI * pow(ly->r[v], 2) * ly->dh) / global_N;
            e[4] += rho * V * qr;
            e[5] += rho * V * qa;
            e[6] = e[6] > diff ? e[6] : diff;
            e[7] = e[7] > fabs(qa) ? e[7] : fabs(qa);
            e[8] = e[8] < qa ? qa : e[8];
            e[9] = e[9] < qa ? e[9] : qa;
        }
    }
    out[0] = e[0] / e[1];
    out[1] = sqrt(e[2] / e[3]);
    out[2] = e[6] / e[7];
    out[3] = e[4];
    out[4] = e[5];
    out[5] = e[9];
    out[6] = e[8];
}
